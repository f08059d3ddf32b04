// Standard field names of the elastic RSG model (reference
// modules/PanNDE/include/StandardNames.hpp:64-167).  The strings are the data format the
// solver looks fields up by, so they are the reference's strings.
#pragma once

#include <string>

namespace PanNDE {

struct ElasticMaterialNames {
  // CIJ[i][j] names the Voigt stiffness entry; the lower triangle names the upper entry.
  std::string CIJ[6][6] = {{"C11", "C12", "C13", "C14", "C15", "C16"},
                           {"C12", "C22", "C23", "C24", "C25", "C26"},
                           {"C13", "C23", "C33", "C34", "C35", "C36"},
                           {"C14", "C24", "C34", "C44", "C45", "C46"},
                           {"C15", "C25", "C35", "C45", "C55", "C56"},
                           {"C16", "C26", "C36", "C46", "C56", "C66"}};
  std::string density = "density";
};

struct ElasticStateNames {
  std::string V[3] = {"Vx", "Vy", "Vz"};
  std::string S[3][3] = {{"Sxx", "Sxy", "Sxz"}, {"Sxy", "Syy", "Syz"}, {"Sxz", "Syz", "Szz"}};
  // Voigt order: xx, yy, zz, yz, xz, xy
  std::string SI[6] = {"Sxx", "Syy", "Szz", "Syz", "Sxz", "Sxy"};
};

struct TimeDomainMetadataNames {
  std::string dt = "dt";
  std::string write_times = "write_times";
};

}  // namespace PanNDE
