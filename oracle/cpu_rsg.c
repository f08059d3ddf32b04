/*
 * cpu_rsg.c -- CPU oracle for PanNDE's rotated-staggered-grid (RSG) elastodynamic step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (pannde_amd/csrc, libpannde_rsg.so) never links or calls it.
 *
 * It restates, loop for loop and in the same floating-point operation order, the
 * reference solver (paths relative to the reference tree):
 *   modules/HostSolver/include/HostRSGTDUT.hpp   solve 162-169, updateTsAtCell 278-293,
 *                                               updateVsAtNode 338-362, evalXds 387-402,
 *                                               cellcenter 419-428
 *   modules/HostSolver/include/HostUTModel.hpp   updateTsAtCell + evalMoments 270-311
 *   modules/HostSolver/include/internal/RightHex.hpp  DD?_cc 100-141, DD?_nc 151-192,
 *                                               ncDifferencer 225-231, ccDifferencer 241-247,
 *                                               cellPrep 253-264, nodePrep 272-309
 *   modules/HostData/include/HexMesh.hpp         buildLinks 300-313 (slot = vertex ^ 6)
 *   modules/HostData/include/HostField.hpp       atCell 121-133, atNode 142-154 (/8 rule)
 *   modules/HostData/include/HostUnivariate.hpp  evalAt 109-121, interpolate 132-142
 *   modules/Stubs/include/HostNormalXDZ.hpp      evalAt 117-123, checkAt 140-145
 *
 * Pinning (what checks each part; see tests/test_oracle.py):
 *   links / nodePrep / cellPrep / DD operators  -> RightHex KATs (HostRSGTDUTTests.hpp:158-256),
 *                                                  closed forms on linear fields, golden
 *                                                  vectors from the compiled reference
 *   field averaging                              -> FieldTests.hpp:216-224 rule, golden vectors
 *   univariate / disc window                     -> HostUnivariateTests.hpp:52-61,
 *                                                  HostNormalXDZTests.hpp:58-68
 *   full multi-step solve (both model forms)     -> bitwise against the compiled reference
 *                                                  (oracle/_ref, tests/golden fixtures),
 *                                                  plus invariants (rigid translation,
 *                                                  linearity, uniform strain rate)
 *   driver step schedule (orc_schedule)          -> exact-arithmetic cases (power-of-two dt:
 *                                                  exactly `interval` steps per write); the
 *                                                  time it compares is the reference's,
 *                                                  bitwise, in every golden fixture
 *
 * Build with -O2 -ffp-contract=off (oracle/Makefile) so no multiply-add is fused.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Planted-mistake switch for the tests that prove the checks can fail: 0 = faithful
 * restatement; 1..5 = one plausible mistake each (see orc_set_mutation).  Never used
 * outside tests/. */
static int g_mut = 0;
void orc_set_mutation(int m) { g_mut = m; }

#define ORC_NODE 0
#define ORC_CELL 1
#define ORC_CONSTANT 2

typedef struct {
  const double* data;
  int32_t type; /* ORC_NODE / ORC_CELL / ORC_CONSTANT */
} orc_field;

/* A circular normal source: HostNormalXDZ(center, radius, ds, HostUnivariate(t, v)). */
typedef struct {
  int32_t component; /* force: direction 0..2; moment: Voigt index 0..5 */
  double center[3];
  double radius;
  double ds;
  int32_t n_signal;
  const double* sig_t;
  const double* sig_v;
} orc_disc;

typedef struct {
  int32_t n_nodes, n_cells;
  const double *x, *y, *z;
  const int32_t* cells; /* [n_cells][8], CGNS vertex order */
  orc_field rho;
  orc_field c[21]; /* C11,C12,..,C16,C22,..,C26,C33,..,C66 */
  double dt;
  int32_t n_force;
  const orc_disc* force; /* listed in transducer order */
  int32_t n_moment;
  const orc_disc* moment; /* listed in excitation order */
  int32_t moment_form; /* 1: HostUTModel (stress source term present), 0: HostRSGTDUT */
} orc_problem;

/* ---------------------------------------------------------------- mesh prep */

/* HexMesh::buildLinks (HexMesh.hpp:300-313): vertex v of cell k puts k in slot v^6. */
void orc_build_links(int32_t nn, int32_t nc, const int32_t* cells, int32_t* node_cells) {
  for (int64_t i = 0; i < (int64_t)nn * 8; i++) node_cells[i] = -1;
  for (int32_t k = 0; k < nc; k++)
    for (int v = 0; v < 8; v++)
      node_cells[(int64_t)cells[(int64_t)k * 8 + v] * 8 + (g_mut == 1 ? v : (v ^ 6))] = k;
}

static inline double coord(const double* x, const double* y, const double* z, int32_t n, int d) {
  return d == 0 ? x[n] : (d == 1 ? y[n] : z[n]);
}

/* RightHex::cellPrep (RightHex.hpp:253-264). */
void orc_cell_prep(int32_t nc, const double* x, const double* y, const double* z,
                   const int32_t* cells, double* dxc, double* dyc, double* dzc) {
  for (int32_t c = 0; c < nc; c++) {
    int32_t a = cells[(int64_t)c * 8 + 0], b = cells[(int64_t)c * 8 + 6];
    dxc[c] = x[b] - x[a];
    dyc[c] = y[b] - y[a];
    dzc[c] = z[b] - z[a];
  }
}

/* RightHex::nodePrep (RightHex.hpp:272-309).  Returns 0, or 1/2/3 for the DX/DY/DZ
 * logic_error the reference throws when a spacing is not positive. */
int orc_node_prep(int32_t nn, const double* x, const double* y, const double* z,
                  const int32_t* cells, const int32_t* node_cells, double* dxn, double* dyn,
                  double* dzn) {
  for (int32_t n = 0; n < nn; n++) {
    double pn[3], lo[3], hi[3];
    for (int d = 0; d < 3; d++) pn[d] = lo[d] = hi[d] = coord(x, y, z, n, d);
    for (int kb = 0; kb < 8; kb++) {
      int32_t c = node_cells[(int64_t)n * 8 + kb];
      if (c == -1) continue;
      for (int kbb = 0; kbb < 8; kbb++) {
        int32_t m = cells[(int64_t)c * 8 + kbb];
        for (int d = 0; d < 3; d++) {
          double p = coord(x, y, z, m, d);
          lo[d] = (p < lo[d]) ? p : lo[d]; /* std::min(a,b) returns a unless b<a */
          hi[d] = (hi[d] < p) ? p : hi[d]; /* std::max(a,b) returns a unless a<b */
        }
      }
    }
    for (int d = 0; d < 3; d++) {
      if (lo[d] == pn[d]) lo[d] = 2. * pn[d] - hi[d];
      if (hi[d] == pn[d]) hi[d] = 2. * pn[d] - lo[d];
    }
    dxn[n] = 0.5 * (hi[0] - lo[0]);
    if (0 >= dxn[n]) return 1;
    dyn[n] = 0.5 * (hi[1] - lo[1]);
    if (0 >= dyn[n]) return 2;
    dzn[n] = 0.5 * (hi[2] - lo[2]);
    if (0 >= dzn[n]) return 3;
  }
  return 0;
}

/* HostField::atNode (HostField.hpp:142-154): NODE fields direct, otherwise the sum over
 * present connected cells divided by 8 (missing cells count as zero). */
double orc_field_at_node(orc_field f, const int32_t* node_cells, int32_t n) {
  if (f.type == ORC_NODE) return f.data[n];
  double v = 0.0;
  int cnt = 0;
  for (int k = 0; k < 8; k++) {
    int32_t c = node_cells[(int64_t)n * 8 + k];
    v += (-1 == c) ? 0. : (f.type == ORC_CONSTANT ? f.data[0] : f.data[c]);
    cnt += (-1 != c);
  }
  return g_mut == 2 ? v / cnt : v / 8.0;
}

/* HostField::atCell (HostField.hpp:121-133): NODE fields are averaged over the 8 vertices. */
double orc_field_at_cell(orc_field f, const int32_t* cells, int32_t c) {
  if (f.type == ORC_NODE) {
    double v = 0.0;
    for (int k = 0; k < 8; k++) {
      int32_t m = cells[(int64_t)c * 8 + k];
      v += (-1 == m) ? 0. : f.data[m];
    }
    return v / 8.0;
  }
  return f.type == ORC_CONSTANT ? f.data[0] : f.data[c];
}

/* ---------------------------------------------------------------- sources */

/* HostUnivariate stores (t, v) in a std::map: keys ascend and a repeated key keeps its
 * first value.  This builds that key order into t_out/v_out and returns the count. */
int32_t orc_signal_sorted(int32_t n, const double* t, const double* v, double* t_out,
                          double* v_out) {
  int32_t m = 0;
  for (int32_t i = 0; i < n; i++) {
    int32_t pos = 0, dup = 0;
    while (pos < m && t_out[pos] < t[i]) pos++;
    if (pos < m && !(t[i] < t_out[pos]) && !(t_out[pos] < t[i])) dup = 1;
    if (dup) continue;
    memmove(t_out + pos + 1, t_out + pos, sizeof(double) * (size_t)(m - pos));
    memmove(v_out + pos + 1, v_out + pos, sizeof(double) * (size_t)(m - pos));
    t_out[pos] = t[i];
    v_out[pos] = v[i];
    m++;
  }
  return m;
}

/* HostUnivariate::evalAt / interpolate on a key-sorted table.  t below the first key is
 * undefined behaviour in the reference (std::prev(begin)); it is reported as NaN here. */
double orc_signal_eval(int32_t m, const double* ts, const double* vs, double t) {
  if (!(t < ts[m - 1])) return vs[m - 1];
  for (int32_t i = 0; i < m; i++) {
    if (t < ts[i]) {
      if (i == 0) return NAN;
      double Dt = ts[i] - ts[i - 1];
      double Dv = vs[i] - vs[i - 1];
      return (Dv / Dt) * (t - ts[i - 1]) + vs[i - 1];
    }
  }
  return 0.;
}

typedef struct {
  const orc_disc* d;
  int32_t m;
  double* ts;
  double* vs;
} orc_disc_rt;

static inline double sqr(double a) { return a * a; }

/* HostNormalXDZ::checkAt (HostNormalXDZ.hpp:140-145). */
int orc_disc_check(const double center[3], double radius, double ds, const double p[3]) {
  double res = 0.0;
  for (int k = 0; k < 3; k++) res += sqr(p[k] - center[k]);
  double dist = sqrt(res);
  double distz = fabs(p[2] - center[2]);
  return (distz <= 0.75 * ds) * (dist < radius) * (dist >= 0.);
}

/* HostNormalXDZ::evalAt (HostNormalXDZ.hpp:117-123); args = {x, y, z, t}. */
static double disc_eval(const orc_disc_rt* r, const double args[4]) {
  const orc_disc* d = r->d;
  double res = 0.0;
  for (int k = 0; k < 3; k++) res += sqr(args[k] - d->center[k]);
  double dist = sqrt(res);
  double scale = orc_signal_eval(r->m, r->ts, r->vs, args[3]);
  double distz = fabs(args[2] - d->center[2]);
  double window = (distz <= 0.75 * d->ds) * (dist < d->radius) * (dist >= 0.);
  return scale * window;
}

/* HostRSGTDUT::cellcenter (HostRSGTDUT.hpp:419-428). */
void orc_cell_center(const double* x, const double* y, const double* z, const int32_t* cells,
                     int32_t c, double cc[3]) {
  cc[0] = 0.;
  cc[1] = 0.;
  cc[2] = 0.;
  for (int kb = 0; kb < 8; kb++) {
    int32_t m = cells[(int64_t)c * 8 + kb];
    cc[0] += 0.125 * x[m];
    cc[1] += 0.125 * y[m];
    cc[2] += 0.125 * z[m];
  }
}

/* ---------------------------------------------------------------- the step */

static inline int cidx(int i, int j) {
  int a = i < j ? i : j, b = i < j ? j : i;
  return a * 6 - a * (a - 1) / 2 + (b - a);
}

/* Slot patterns of RightHex (RightHex.hpp:100-192): forward / backward vertices (or
 * connected-cell slots) for d/dx, d/dy, d/dz. */
static const int FWD[3][4] = {{1, 2, 5, 6}, {2, 3, 6, 7}, {4, 5, 6, 7}};
static const int BWD[3][4] = {{0, 3, 4, 7}, {0, 1, 4, 5}, {0, 1, 2, 3}};

typedef struct {
  const orc_problem* p;
  int32_t* node_cells;
  double *dxn[3], *dxc[3];
  orc_disc_rt* force;
  orc_disc_rt* moment;
} orc_ctx;

/* RightHex::DD?_nc through ncDifferencer: cell field -> node. */
static inline double dd_nc(const orc_ctx* cx, int axis, int32_t n, const double* f) {
  const int32_t* box = cx->node_cells + (int64_t)n * 8;
  double value = 0.0;
  for (int k = 0; k < 4; k++) {
    int32_t a = box[FWD[axis][k]], b = box[BWD[axis][k]];
    value += (((-1 == a) ? 0. : f[a]) - ((-1 == b) ? 0. : f[b]));
  }
  return (value * 0.25 / cx->dxn[axis][n]);
}

/* RightHex::DD?_cc through ccDifferencer: node field -> cell. */
static inline double dd_cc(const orc_ctx* cx, int axis, int32_t c, const double* f) {
  const int32_t* box = cx->p->cells + (int64_t)c * 8;
  double value = 0.0;
  for (int k = 0; k < 4; k++) value += (f[box[FWD[axis][k]]] - f[box[BWD[axis][k]]]);
  if (g_mut == 5 && axis == 1) value = -value;
  return (value * 0.25 / cx->dxc[axis][c]);
}

/* HostRSGTDUT::evalXds(dir, n, time) (HostRSGTDUT.hpp:387-402). */
static double eval_xds(const orc_ctx* cx, int dir, int32_t n, double time) {
  const orc_problem* p = cx->p;
  double result = 0.;
  int any = 0;
  for (int i = 0; i < p->n_force; i++) any |= (p->force[i].component == dir);
  if (!any) return result; /* every transducer is nullptr in this direction: adds 0. */
  for (int kb = 0; kb < 8; kb++) {
    int32_t c = cx->node_cells[(int64_t)n * 8 + kb];
    if (-1 == c) continue;
    double args[4];
    orc_cell_center(p->x, p->y, p->z, p->cells, c, args);
    args[3] = time;
    for (int i = 0; i < p->n_force; i++)
      if (p->force[i].component == dir)
        result += (g_mut == 4 ? 1.0 : 0.125) * (disc_eval(&cx->force[i], args));
  }
  return result;
}

/* HostUTModel::evalMoments(component, c, time) (HostUTModel.hpp:296-311). */
static double eval_moments(const orc_ctx* cx, int comp, int32_t c, double time) {
  const orc_problem* p = cx->p;
  double result = 0.;
  int any = 0;
  for (int i = 0; i < p->n_moment; i++) any |= (p->moment[i].component == comp);
  if (!any) return result;
  for (int kb = 0; kb < 8; kb++) {
    int32_t m = p->cells[(int64_t)c * 8 + kb];
    double args[4] = {p->x[m], p->y[m], p->z[m], time};
    for (int i = 0; i < p->n_moment; i++)
      if (p->moment[i].component == comp) result += 0.125 * (disc_eval(&cx->moment[i], args));
  }
  return result;
}

static int ctx_init(orc_ctx* cx, const orc_problem* p) {
  memset(cx, 0, sizeof(*cx));
  cx->p = p;
  int32_t nn = p->n_nodes, nc = p->n_cells;
  cx->node_cells = (int32_t*)malloc(sizeof(int32_t) * 8 * (size_t)nn);
  for (int d = 0; d < 3; d++) {
    cx->dxn[d] = (double*)malloc(sizeof(double) * (size_t)nn);
    cx->dxc[d] = (double*)malloc(sizeof(double) * (size_t)nc);
  }
  orc_build_links(nn, nc, p->cells, cx->node_cells);
  int err = orc_node_prep(nn, p->x, p->y, p->z, p->cells, cx->node_cells, cx->dxn[0], cx->dxn[1],
                          cx->dxn[2]);
  orc_cell_prep(nc, p->x, p->y, p->z, p->cells, cx->dxc[0], cx->dxc[1], cx->dxc[2]);
  cx->force = (orc_disc_rt*)calloc((size_t)p->n_force + 1, sizeof(orc_disc_rt));
  cx->moment = (orc_disc_rt*)calloc((size_t)p->n_moment + 1, sizeof(orc_disc_rt));
  for (int pass = 0; pass < 2; pass++) {
    int cnt = pass ? p->n_moment : p->n_force;
    const orc_disc* src = pass ? p->moment : p->force;
    orc_disc_rt* dst = pass ? cx->moment : cx->force;
    for (int i = 0; i < cnt; i++) {
      dst[i].d = &src[i];
      dst[i].ts = (double*)malloc(sizeof(double) * (size_t)src[i].n_signal);
      dst[i].vs = (double*)malloc(sizeof(double) * (size_t)src[i].n_signal);
      dst[i].m = orc_signal_sorted(src[i].n_signal, src[i].sig_t, src[i].sig_v, dst[i].ts, dst[i].vs);
    }
  }
  return err;
}

static void ctx_free(orc_ctx* cx) {
  free(cx->node_cells);
  for (int d = 0; d < 3; d++) {
    free(cx->dxn[d]);
    free(cx->dxc[d]);
  }
  for (int i = 0; i < cx->p->n_force; i++) {
    free(cx->force[i].ts);
    free(cx->force[i].vs);
  }
  for (int i = 0; i < cx->p->n_moment; i++) {
    free(cx->moment[i].ts);
    free(cx->moment[i].vs);
  }
  free(cx->force);
  free(cx->moment);
}

/* Runs nsteps of Model::solve(1) on state V ([3][n_nodes]: Vx,Vy,Vz) and S ([6][n_cells]:
 * Sxx,Syy,Szz,Syz,Sxz,Sxy), starting at *time and advancing it by time += dt per step.
 * After every step the velocities at the n_sens sensor nodes are written to
 * traces[step][sensor][3] (if traces is not NULL).  Returns 0, or the nodePrep error. */
int orc_run(const orc_problem* p, int32_t nsteps, double* V, double* S, double* time,
            int32_t n_sens, const int32_t* sens, double* traces) {
  orc_ctx cx;
  int err = ctx_init(&cx, p);
  if (err) {
    ctx_free(&cx);
    return err;
  }
  const int32_t nn = p->n_nodes, nc = p->n_cells;
  double* Vc[3] = {V, V + nn, V + 2 * (int64_t)nn};
  double* Sc[6];
  for (int k = 0; k < 6; k++) Sc[k] = S + (int64_t)k * nc;
  const double dt = p->dt;
  double t = *time;
  for (int32_t step = 0; step < nsteps; step++) {
    /* updateVs -> updateVsAtNode (HostRSGTDUT.hpp:338-362); Voigt S[0..5] =
     * Sxx,Syy,Szz,Syz,Sxz,Sxy. */
    for (int32_t n = 0; n < nn; n++) {
      Vc[0][n] += (dt / (orc_field_at_node(p->rho, cx.node_cells, n))) *
                  (dd_nc(&cx, 0, n, Sc[0]) + dd_nc(&cx, 1, n, Sc[g_mut == 3 ? 3 : 5]) +
                   dd_nc(&cx, 2, n, Sc[4]) + eval_xds(&cx, 0, n, t));
      Vc[1][n] += (dt / (orc_field_at_node(p->rho, cx.node_cells, n))) *
                  (dd_nc(&cx, 0, n, Sc[5]) + dd_nc(&cx, 1, n, Sc[1]) + dd_nc(&cx, 2, n, Sc[3]) +
                   eval_xds(&cx, 1, n, t));
      Vc[2][n] += (dt / (orc_field_at_node(p->rho, cx.node_cells, n))) *
                  (dd_nc(&cx, 0, n, Sc[4]) + dd_nc(&cx, 1, n, Sc[3]) + dd_nc(&cx, 2, n, Sc[2]) +
                   eval_xds(&cx, 2, n, t));
    }
    /* updateTs -> updateTsAtCell (HostRSGTDUT.hpp:278-293, HostUTModel.hpp:270-286). */
    for (int32_t c = 0; c < nc; c++) {
      double dv[3][3];
      for (int k1 = 0; k1 < 3; k1++) {
        dv[0][k1] = dd_cc(&cx, 0, c, Vc[k1]);
        dv[1][k1] = dd_cc(&cx, 1, c, Vc[k1]);
        dv[2][k1] = dd_cc(&cx, 2, c, Vc[k1]);
      }

      for (int ks = 0; ks < 6; ks++) {
        double c0 = orc_field_at_cell(p->c[cidx(ks, 0)], p->cells, c);
        double c1 = orc_field_at_cell(p->c[cidx(ks, 1)], p->cells, c);
        double c2 = orc_field_at_cell(p->c[cidx(ks, 2)], p->cells, c);
        double c3 = orc_field_at_cell(p->c[cidx(ks, 3)], p->cells, c);
        double c4 = orc_field_at_cell(p->c[cidx(ks, 4)], p->cells, c);
        double c5 = orc_field_at_cell(p->c[cidx(ks, 5)], p->cells, c);
        if (p->moment_form) {
          Sc[ks][c] += dt * (c0 * dv[0][0] + c1 * dv[1][1] + c2 * dv[2][2] +
                             c3 * (dv[1][2] + dv[2][1]) + c4 * (dv[0][2] + dv[2][0]) +
                             c5 * (dv[0][1] + dv[1][0]) + eval_moments(&cx, ks, c, t));
        } else {
          Sc[ks][c] += dt * (c0 * dv[0][0] + c1 * dv[1][1] + c2 * dv[2][2] +
                             c3 * (dv[1][2] + dv[2][1]) + c4 * (dv[0][2] + dv[2][0]) +
                             c5 * (dv[0][1] + dv[1][0]));
        }
      }
    }
    t += dt;
    if (traces)
      for (int32_t s = 0; s < n_sens; s++)
        for (int d = 0; d < 3; d++)
          traces[((int64_t)step * n_sens + s) * 3 + d] = Vc[d][sens[s]];
  }
  *time = t;
  ctx_free(&cx);
  return 0;
}

/* RightHex helpers exposed for the known-answer tests: D of a node field at cell c and
 * of a cell field at node n, on an axis 0..2, plus the node/cell spacings. */
int orc_dd_probe(int32_t nn, int32_t nc, const double* x, const double* y, const double* z,
                 const int32_t* cells, const double* node_field, const double* cell_field,
                 double* dd_cc_out /* [nc][3] */, double* dd_nc_out /* [nn][3] */,
                 double* dn_out /* [nn][3] */, double* dc_out /* [nc][3] */) {
  orc_problem p;
  memset(&p, 0, sizeof(p));
  p.n_nodes = nn;
  p.n_cells = nc;
  p.x = x;
  p.y = y;
  p.z = z;
  p.cells = cells;
  orc_ctx cx;
  int err = ctx_init(&cx, &p);
  if (!err) {
    for (int32_t c = 0; c < nc; c++)
      for (int a = 0; a < 3; a++) {
        dd_cc_out[(int64_t)c * 3 + a] = dd_cc(&cx, a, c, node_field);
        dc_out[(int64_t)c * 3 + a] = cx.dxc[a][c];
      }
    for (int32_t n = 0; n < nn; n++)
      for (int a = 0; a < 3; a++) {
        dd_nc_out[(int64_t)n * 3 + a] = dd_nc(&cx, a, n, cell_field);
        dn_out[(int64_t)n * 3 + a] = cx.dxn[a][n];
      }
  }
  ctx_free(&cx);
  return err;
}

/* Driver step schedule (TransducerModel.cpp:63-83): for every write time, step while
 * time < write_time.  steps_out[k] = steps taken before write k.  Returns total steps. */
int64_t orc_schedule(double dt, int32_t n_write, const double* write_times, int64_t* steps_out) {
  double t = 0;
  int64_t total = 0;
  for (int32_t k = 0; k < n_write; k++) {
    int64_t s = 0;
    while (t < write_times[k]) {
      t += dt;
      s++;
    }
    steps_out[k] = s;
    total += s;
  }
  return total;
}
