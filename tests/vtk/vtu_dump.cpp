// vtu_dump.cpp -- prints what pannde_amd/csrc/vtk_job.hpp reads from a .vtu file, as text, so a
// test can compare it with pannde_amd.vtkio.read_vtu.  TEST INFRASTRUCTURE ONLY.
//   points N / cells M, then "points" and "cells" lines, then one line per array:
//   <P|C|F> <name> <components> <values...>  (PointData, CellData, FieldData); errors go to
//   stderr with exit status 1.
#include <cstdio>

#include "../../pannde_amd/csrc/vtk_job.hpp"

int main(int argc, char** argv) {
  if (argc != 2) return 2;
  try {
    const vtkjob::Vtu v = vtkjob::read_vtu(argv[1]);
    std::printf("points %d\ncells %d\npoints", v.npts, v.ncells);
    for (int i = 0; i < v.npts; i++) std::printf(" %.17g %.17g %.17g", v.x[i], v.y[i], v.z[i]);
    std::printf("\ncells");
    for (int32_t c : v.cells) std::printf(" %d", c);
    std::printf("\n");
    const std::pair<char, const std::map<std::string, vtkjob::Array>*> secs[3] = {
        {'P', &v.point_data}, {'C', &v.cell_data}, {'F', &v.field_data}};
    for (auto& s : secs)
      for (auto& kv : *s.second) {
        std::printf("%c %s %d", s.first, kv.first.c_str(), kv.second.ncomp);
        for (double x : kv.second.v) std::printf(" %.17g", x);
        std::printf("\n");
      }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
