// CPU oracle: a C++ restatement of GridSim's CPU sparse Newton-Raphson load flow (TEST INFRASTRUCTURE ONLY).
//
// Used by tests/, the smoke check and the benchmark's CPU-baseline leg as the checker; the product never links it.
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off; no FMA anywhere, as the reference forbids it,
// MultiCaseSparseLu.cs:10-16).
//
// Reference files restated here (paths relative to the reference root):
//   src/GridSim.Core/PowerFlow/SparseNewtonRaphsonPowerFlow.cs  (BuildY :382-402, ComputeInjections :269-289,
//       RunNewton :168-265, BuildJacobianStructure :296-332, RefillJacobian :342-380)
//   src/GridSim.Core/PowerFlow/NewtonShell.cs  (SolveStep :74-120, ApplyStep :128-136, DivergedBeyondLimit :177-178,
//       EquationPositions :327-336, BuildBusSpec :433-486, AssembleResult :649-697)
//   src/GridSim.Core/PowerFlow/AdmittanceStamp.cs:35-75, CpuSparseLuSolver.cs:30-48
//   src/GridSim.Core/Numerics/Sparse.cs  (CsrMatrix.Builder :33-124, ComplexCsr.Builder :145-222, Ilu0 :230-300,
//       BiCgStab :302-405)
//   src/GridSim.Core/Numerics/SparseLu.cs  (ToCsc :63-91, ToCscKeepingZeros :97-117)
//   src/GridSim.Core/Numerics/SparseLuFactorization.cs  (Factor :69-195, Refactor :246-306, Solve :338-371,
//       SpSolve :753-772, Reach :775-786, Dfs :789-821)
//   src/GridSim.Core/Numerics/MinimumDegreeOrdering.cs:23-136
//   src/GridSim.Core/Numerics/NumericTolerances.cs  (RelativePivotFloor :36, row floors :99-115, MaxAbs :165,
//       Ilu0PivotGuard :226-234, UnderflowGuard :242, Krylov cap :256-260)
//   src/GridSim.Core/Analysis/ContingencyAnalysis.cs  (ScreenBranchesParallel :323-347, WithoutMany :1839-1845)
//   src/GridSim.Core/PowerFlow/NewtonShell.cs  (Stalled :162,200-206, SolveWithRetry :227-268,
//       TrySeedDcAngles :282-299, MaxBranchAngleDeg :311-320), SparseNewtonRaphsonPowerFlow.cs (SolveAcAttempt :71-165)
//   src/GridSim.Core/Analysis/DcPowerFlow.cs :78-86, DcDemand.cs (LegacyView :38-50),
//       DistributionFactors.cs (Prepare :143-248, IslandsCore :261-360, DcSolve / InjectionRhs :454-500)
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "fdlibm.h"

namespace {

using std::vector;
constexpr double kPi = 3.141592653589793;  // Math.PI
constexpr double kRelativePivotFloor = 1e-12;
constexpr double kMaxUndampedStep = 0.4;
constexpr double kVmFloorPu = 0.01;
constexpr double kUnderflowGuard = 1e-300;

// .NET Math.Max(double, double): NaN if either argument is NaN.
inline double net_max(double a, double b) {
  if (std::isnan(a) || std::isnan(b)) return std::numeric_limits<double>::quiet_NaN();
  return a > b ? a : b;
}

// ---------------------------------------------------------------- System.Numerics.Complex arithmetic
struct Cx { double re, im; };
inline Cx cx_add(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
inline Cx cx_neg(Cx a) { return {-a.re, -a.im}; }
inline Cx cx_conj(Cx a) { return {a.re, -a.im}; }
// Complex / Complex: Smith's formula (the .NET BCL operator).
inline Cx cx_div(Cx l, Cx r) {
  double a = l.re, b = l.im, c = r.re, d = r.im;
  if (std::fabs(d) < std::fabs(c)) {
    double doc = d / c;
    return {(a + b * doc) / (c + d * doc), (b - a * doc) / (c + d * doc)};
  }
  double cod = c / d;
  return {(b + a * cod) / (d + c * cod), (-a + b * cod) / (d + c * cod)};
}
// Complex / double (finite operands: component-wise division).
inline Cx cx_div_real(Cx l, double r) {
  if (r == 0) return {std::nan(""), std::nan("")};
  return {l.re / r, l.im / r};
}

// Complex * Complex: (ac - bd) + (bc + ad) i; Complex * double: component-wise.
inline Cx cx_mul(Cx l, Cx r) { return {l.re * r.re - l.im * r.im, l.im * r.re + l.re * r.im}; }
inline Cx cx_scale(Cx l, double d) { return {l.re * d, l.im * d}; }
// Complex.Magnitude: the scaled hypotenuse |large| sqrt(1 + (small / large)^2).
inline double cx_abs(Cx z) {
  double a = std::fabs(z.re), b = std::fabs(z.im);
  double small = a < b ? a : b, large = a < b ? b : a;
  if (small == 0.0) return large;
  if (std::isinf(large) && !std::isnan(small)) return std::numeric_limits<double>::infinity();
  double ratio = small / large;
  return large * std::sqrt(1.0 + ratio * ratio);
}

struct Stamp { Cx yff, ytt, yft, ytf; };

// AdmittanceStamp.Branch + FromAdmittance (AdmittanceStamp.cs:41-75) at h = 1, rFactor = 1.
Stamp branch_stamp(double r, double x, double b, double tap_ratio, double shift_deg) {
  Cx ys = cx_div(Cx{1.0, 0.0}, Cx{r * 1.0, x * 1.0});
  Cx bc = {0.0, (b * 1.0) / 2.0};
  double a = tap_ratio == 0 ? 1.0 : tap_ratio;
  double shift = shift_deg * kPi / 180.0;
  double s, c;
  orc_fdlibm::sincos(shift, &s, &c);
  Cx tap = {a * c, a * s};  // DetMath.FromPolar
  Stamp st;
  st.yff = cx_div_real(cx_add(ys, bc), a * a);
  st.ytt = cx_add(ys, bc);
  st.yft = cx_div(cx_neg(ys), cx_conj(tap));
  st.ytf = cx_div(cx_neg(ys), tap);
  return st;
}

// ---------------------------------------------------------------- ComplexCsr.Builder (Sparse.cs:145-222)
struct Ycsr {
  int n = 0;
  vector<int> rp, ci, diag;
  vector<double> g, b;
};

struct YBuilder {
  int n;
  vector<int> r, c;
  vector<double> g, b;
  explicit YBuilder(int n_) : n(n_) {}
  void add(int row, int col, double gv, double bv) {
    if (gv == 0.0 && bv == 0.0) return;
    r.push_back(row); c.push_back(col); g.push_back(gv); b.push_back(bv);
  }
  Ycsr build() const {
    int nnz = (int)g.size();
    vector<int> row_start(n + 1, 0);
    for (int t = 0; t < nnz; t++) row_start[r[t] + 1]++;
    for (int i = 0; i < n; i++) row_start[i + 1] += row_start[i];
    vector<int> col(nnz), fill(row_start.begin(), row_start.end() - 1);
    vector<double> gv(nnz), bv(nnz);
    for (int t = 0; t < nnz; t++) { int p = fill[r[t]]++; col[p] = c[t]; gv[p] = g[t]; bv[p] = b[t]; }
    Ycsr y;
    y.n = n;
    y.rp.assign(n + 1, 0);
    y.diag.assign(n, -1);
    vector<int> order;
    for (int i = 0; i < n; i++) {
      int s = row_start[i], e = row_start[i + 1], len = e - s;
      order.resize(len);
      for (int k = 0; k < len; k++) order[k] = s + k;
      // stable by (column, insertion index)
      std::stable_sort(order.begin(), order.end(), [&](int x, int z) { return col[x] < col[z]; });
      y.rp[i] = (int)y.ci.size();
      int prev = std::numeric_limits<int>::min();
      bool diag_written = false;
      for (int k = 0; k < len; k++) {
        int cc = col[order[k]];
        double gg = gv[order[k]], bb = bv[order[k]];
        if (cc == prev) { y.g.back() += gg; y.b.back() += bb; continue; }
        if (!diag_written && prev < i && cc > i) {
          y.diag[i] = (int)y.ci.size(); y.ci.push_back(i); y.g.push_back(0); y.b.push_back(0); diag_written = true;
        }
        y.ci.push_back(cc); y.g.push_back(0.0 + gg); y.b.push_back(0.0 + bb); prev = cc;
        if (cc == i) { y.diag[i] = (int)y.ci.size() - 1; diag_written = true; }
      }
      if (!diag_written) { y.diag[i] = (int)y.ci.size(); y.ci.push_back(i); y.g.push_back(0); y.b.push_back(0); }
    }
    y.rp[n] = (int)y.ci.size();
    return y;
  }
};

// BuildY (SparseNewtonRaphsonPowerFlow.cs:382-402); `skip` removes one branch (WithoutMany keeps model order).
Ycsr build_y(const orc_network& net, int skip) {
  YBuilder bld(net.n_bus);
  for (int k = 0; k < net.n_branch; k++) {
    if (k == skip || !net.br_in_service[k]) continue;
    int f = net.br_from[k], t = net.br_to[k];
    Stamp s = branch_stamp(net.br_r[k], net.br_x[k], net.br_b[k], net.br_tap[k], net.br_shift_deg[k]);
    bld.add(f, f, s.yff.re, s.yff.im);
    bld.add(t, t, s.ytt.re, s.ytt.im);
    bld.add(f, t, s.yft.re, s.yft.im);
    bld.add(t, f, s.ytf.re, s.ytf.im);
  }
  for (int i = 0; i < net.n_bus; i++) {
    if (net.gs[i] == 0 && net.bs[i] == 0) continue;
    bld.add(i, i, net.gs[i] / net.base_mva, net.bs[i] / net.base_mva);
  }
  return bld.build();
}

// ---------------------------------------------------------------- CsrMatrix (Sparse.cs:10-124)
struct Csr {
  int n = 0;
  vector<int> rp, ci, diag;
  vector<double> val;
};

struct CsrBuilder {
  int n;
  vector<int> r, c;
  vector<double> v;
  explicit CsrBuilder(int n_) : n(n_) {}
  void add(int row, int col, double value) {
    if (value == 0.0) return;
    r.push_back(row); c.push_back(col); v.push_back(value);
  }
  Csr build() const {
    int nnz_raw = (int)v.size();
    vector<int> count(n + 1, 0);
    for (int t = 0; t < nnz_raw; t++) count[r[t] + 1]++;
    vector<int> row_start(n + 1, 0);
    for (int i = 0; i < n; i++) row_start[i + 1] = row_start[i] + count[i + 1];
    vector<int> col_tmp(nnz_raw);
    vector<double> val_tmp(nnz_raw);
    vector<int> fill(row_start.begin(), row_start.end() - 1);
    for (int t = 0; t < nnz_raw; t++) { int p = fill[r[t]]++; col_tmp[p] = c[t]; val_tmp[p] = v[t]; }
    Csr a;
    a.n = n;
    a.rp.assign(n + 1, 0);
    a.diag.assign(n, -1);
    vector<int> idx;
    for (int i = 0; i < n; i++) {
      int s = row_start[i], e = row_start[i + 1], len = e - s;
      idx.resize(len);
      for (int k = 0; k < len; k++) idx[k] = s + k;
      std::stable_sort(idx.begin(), idx.end(), [&](int x, int z) { return col_tmp[x] < col_tmp[z]; });
      a.rp[i] = (int)a.ci.size();
      int prev = std::numeric_limits<int>::min();
      bool diag_written = false;
      for (int k = 0; k < len; k++) {
        int cc = col_tmp[idx[k]];
        double vv = val_tmp[idx[k]];
        if (cc == prev) { a.val.back() += vv; continue; }
        if (!diag_written && prev < i && cc > i) {
          a.diag[i] = (int)a.ci.size(); a.ci.push_back(i); a.val.push_back(0.0); diag_written = true;
        }
        a.ci.push_back(cc); a.val.push_back(vv); prev = cc;
        if (cc == i) { a.diag[i] = (int)a.ci.size() - 1; diag_written = true; }
      }
      if (!diag_written) { a.diag[i] = (int)a.ci.size(); a.ci.push_back(i); a.val.push_back(0.0); }
    }
    a.rp[n] = (int)a.ci.size();
    return a;
  }
};

void csr_multiply(const Csr& a, const double* x, double* y) {
  for (int i = 0; i < a.n; i++) {
    double s = 0;
    for (int k = a.rp[i]; k < a.rp[i + 1]; k++) s += a.val[k] * x[a.ci[k]];
    y[i] = s;
  }
}

// ---------------------------------------------------------------- bus spec, numbering, injections, Jacobian
struct BusSpec {
  vector<int> type0;
  vector<double> vm, va, pspec, qspec;
  vector<double> qd, vg, qmin, qmax;   // MVAr / pu: the reactive-limit loop's data (sums over a bus's generators)
};

// NewtonShell.BuildBusSpec (NewtonShell.cs:433-486), constant-power buses.
BusSpec build_bus_spec(const orc_network& net, bool flat_start) {
  int n = net.n_bus;
  BusSpec s;
  s.type0.assign(n, ORC_PQ);
  s.vm.assign(n, 0); s.va.assign(n, 0); s.pspec.assign(n, 0); s.qspec.assign(n, 0);
  vector<double> pg(n, 0), qg(n, 0), vg(n, 0);
  vector<char> has_gen(n, 0);
  s.qmin.assign(n, 0);
  s.qmax.assign(n, 0);
  for (int g = 0; g < net.n_gen; g++) {
    int i = net.gen_bus[g];
    pg[i] += net.gen_pg[g];
    qg[i] += net.gen_qg[g];
    s.qmin[i] += net.gen_qmin ? net.gen_qmin[g] : -INFINITY;
    s.qmax[i] += net.gen_qmax ? net.gen_qmax[g] : INFINITY;
    vg[i] = net.gen_vg[g];
    has_gen[i] = 1;
  }
  s.vg = vg;
  s.qd.assign(net.qd, net.qd + n);
  double base = net.base_mva;
  for (int i = 0; i < n; i++) {
    int bt = net.bus_type[i];
    s.type0[i] = (bt == ORC_PV && !has_gen[i]) ? ORC_PQ : bt;
    bool vc = s.type0[i] == ORC_SLACK || s.type0[i] == ORC_PV;
    s.vm[i] = vc && has_gen[i] ? vg[i] : s.type0[i] == ORC_SLACK ? net.vm[i] : flat_start ? 1.0 : net.vm[i];
    s.va[i] = bt == ORC_SLACK ? net.va_deg[i] * kPi / 180.0 : flat_start ? 0.0 : net.va_deg[i] * kPi / 180.0;
    s.pspec[i] = (pg[i] - net.pd[i]) / base;
    s.qspec[i] = (qg[i] - net.qd[i]) / base;
  }
  return s;
}

// NewtonShell.EquationPositions (NewtonShell.cs:327-336).
int equation_positions(const vector<int>& type, vector<int>& ppos, vector<int>& qpos) {
  int n = (int)type.size();
  ppos.assign(n, -1);
  qpos.assign(n, -1);
  int m = 0;
  for (int i = 0; i < n; i++) if (type[i] != ORC_SLACK) ppos[i] = m++;
  for (int i = 0; i < n; i++) if (type[i] == ORC_PQ) qpos[i] = m++;
  return m;
}

// ComputeInjections (SparseNewtonRaphsonPowerFlow.cs:269-289).
void compute_injections(const Ycsr& y, const double* vm, const double* va, double* p, double* q) {
  for (int i = 0; i < y.n; i++) {
    double pi = 0, qi = 0, vai = va[i];
    for (int k = y.rp[i]; k < y.rp[i + 1]; k++) {
      int j = y.ci[k];
      double th = vai - va[j];
      double sn, cs;
      orc_fdlibm::sincos(th, &sn, &cs);
      double vmj = vm[j];
      pi += vmj * (y.g[k] * cs + y.b[k] * sn);
      qi += vmj * (y.g[k] * sn - y.b[k] * cs);
    }
    p[i] = vm[i] * pi;
    q[i] = vm[i] * qi;
  }
}

int locate(const Csr& a, int row, int col) {
  auto b = a.ci.begin() + a.rp[row], e = a.ci.begin() + a.rp[row + 1];
  auto it = std::lower_bound(b, e, col);
  return (it != e && *it == col) ? (int)(it - a.ci.begin()) : -1;
}

// BuildJacobianStructure (SparseNewtonRaphsonPowerFlow.cs:296-332).
Csr build_jacobian_structure(const Ycsr& y, const vector<int>& ppos, const vector<int>& qpos, int m,
                             vector<int>& slots) {
  int n = y.n;
  CsrBuilder jb(m);
  for (int i = 0; i < n; i++) {
    int pi = ppos[i], qi = qpos[i];
    if (pi < 0 && qi < 0) continue;
    for (int k = y.rp[i]; k < y.rp[i + 1]; k++) {
      int j = y.ci[k];
      int pj = ppos[j], qj = qpos[j];
      if (pi >= 0 && pj >= 0) jb.add(pi, pj, 1.0);
      if (pi >= 0 && qj >= 0) jb.add(pi, qj, 1.0);
      if (qi >= 0 && pj >= 0) jb.add(qi, pj, 1.0);
      if (qi >= 0 && qj >= 0) jb.add(qi, qj, 1.0);
    }
  }
  Csr jac = jb.build();
  slots.assign(4 * (size_t)y.rp[n], -1);
  for (int i = 0; i < n; i++) {
    int pi = ppos[i], qi = qpos[i];
    for (int k = y.rp[i]; k < y.rp[i + 1]; k++) {
      int j = y.ci[k];
      int pj = ppos[j], qj = qpos[j];
      size_t b = 4 * (size_t)k;
      slots[b + 0] = pi >= 0 && pj >= 0 ? locate(jac, pi, pj) : -1;
      slots[b + 1] = pi >= 0 && qj >= 0 ? locate(jac, pi, qj) : -1;
      slots[b + 2] = qi >= 0 && pj >= 0 ? locate(jac, qi, pj) : -1;
      slots[b + 3] = qi >= 0 && qj >= 0 ? locate(jac, qi, qj) : -1;
    }
  }
  return jac;
}

// RefillJacobian (SparseNewtonRaphsonPowerFlow.cs:342-380).
void refill_jacobian(Csr& jac, const vector<int>& slots, const Ycsr& y, const double* vm, const double* va,
                     const double* p, const double* q) {
  std::fill(jac.val.begin(), jac.val.end(), 0.0);
  for (int i = 0; i < y.n; i++) {
    for (int k = y.rp[i]; k < y.rp[i + 1]; k++) {
      int j = y.ci[k];
      double gij = y.g[k], bij = y.b[k];
      double h, nn, mm, l;
      if (i == j) {
        h = -q[i] - bij * vm[i] * vm[i];
        nn = p[i] / vm[i] + gij * vm[i];
        mm = p[i] - gij * vm[i] * vm[i];
        l = q[i] / vm[i] - bij * vm[i];
      } else {
        double th = va[i] - va[j];
        double sn, cs;
        orc_fdlibm::sincos(th, &sn, &cs);
        h = vm[i] * vm[j] * (gij * sn - bij * cs);
        nn = vm[i] * (gij * cs + bij * sn);
        mm = -vm[i] * vm[j] * (gij * cs + bij * sn);
        l = vm[i] * (gij * sn - bij * cs);
      }
      size_t b = 4 * (size_t)k;
      int s;
      if ((s = slots[b + 0]) >= 0) jac.val[s] = h;
      if ((s = slots[b + 1]) >= 0) jac.val[s] = nn;
      if ((s = slots[b + 2]) >= 0) jac.val[s] = mm;
      if ((s = slots[b + 3]) >= 0) jac.val[s] = l;
    }
  }
}

// ---------------------------------------------------------------- minimum degree (MinimumDegreeOrdering.cs:23-136)
vector<int> md_order(int n, const int* rowptr, const int* colidx) {
  vector<int> order(n);
  if (n == 0) return order;
  vector<int> cnt(n, 0);
  for (int i = 0; i < n; i++)
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
      int j = colidx[k];
      if (j != i) { cnt[i]++; cnt[j]++; }
    }
  vector<vector<int>> adj(n);
  for (int i = 0; i < n; i++) adj[i].assign(cnt[i], 0);
  vector<int> len(n, 0);
  for (int i = 0; i < n; i++)
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
      int j = colidx[k];
      if (j == i) continue;
      adj[i][len[i]++] = j;
      adj[j][len[j]++] = i;
    }
  vector<int> stamp(n, 0);
  int st = 0;
  for (int i = 0; i < n; i++) {
    st++;
    int w = 0;
    for (int p = 0; p < len[i]; p++) {
      int j = adj[i][p];
      if (stamp[j] != st) { stamp[j] = st; adj[i][w++] = j; }
    }
    len[i] = w;
  }
  vector<int> head(n + 1, -1), next(n), prev(n), deg(n);
  auto insert = [&](int v, int d) {
    deg[v] = d; prev[v] = -1; next[v] = head[d];
    if (head[d] >= 0) prev[head[d]] = v;
    head[d] = v;
  };
  auto remove = [&](int v) {
    int d = deg[v];
    if (prev[v] >= 0) next[prev[v]] = next[v]; else head[d] = next[v];
    if (next[v] >= 0) prev[next[v]] = prev[v];
  };
  for (int i = n - 1; i >= 0; i--) insert(i, len[i]);
  vector<char> dead(n, 0);
  vector<int> nb(n);
  int min_deg = 0;
  for (int s = 0; s < n; s++) {
    while (head[min_deg] < 0) min_deg++;
    int v = head[min_deg];
    remove(v);
    dead[v] = 1;
    order[s] = v;
    int nb_count = 0;
    for (int p = 0; p < len[v]; p++) {
      int u = adj[v][p];
      if (!dead[u]) nb[nb_count++] = u;
    }
    st++;
    for (int t = 0; t < nb_count; t++) stamp[nb[t]] = st;
    int nb_stamp = st;
    for (int t = 0; t < nb_count; t++) {
      int u = nb[t];
      remove(u);
      vector<int>& au = adj[u];
      int w = 0;
      for (int p = 0; p < len[u]; p++) {
        int x = au[p];
        if (x != v && !dead[x] && stamp[x] != nb_stamp) au[w++] = x;
      }
      int need = w + nb_count - 1;
      if ((int)au.size() < need) au.resize(std::max(need, (int)au.size() * 2));
      for (int qq = 0; qq < nb_count; qq++)
        if (nb[qq] != u) au[w++] = nb[qq];
      len[u] = w;
      insert(u, w);
    }
    adj[v].clear();
    len[v] = 0;
    if (min_deg > 0) min_deg--;
  }
  return order;
}

// ---------------------------------------------------------------- CSR -> CSC (SparseLu.cs:63-117)
struct Csc {
  int n = 0;
  vector<int> ap, ai;
  vector<double> ax;
};

Csc to_csc(int n, const int* rowptr, const int* colidx, const double* val, bool keep_zeros) {
  Csc c;
  c.n = n;
  c.ap.assign(n + 1, 0);
  int stored = rowptr[n];
  for (int k = 0; k < stored; k++)
    if (keep_zeros || val[k] != 0.0) c.ap[colidx[k] + 1]++;
  for (int j = 0; j < n; j++) c.ap[j + 1] += c.ap[j];
  int nnz = c.ap[n];
  c.ai.assign(nnz, 0);
  c.ax.assign(nnz, 0.0);
  vector<int> fill(c.ap.begin(), c.ap.end() - 1);
  for (int i = 0; i < n; i++)
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
      double v = val[k];
      if (!keep_zeros && v == 0.0) continue;
      int dst = fill[colidx[k]]++;
      c.ai[dst] = i;
      c.ax[dst] = v;
    }
  return c;
}

// ---------------------------------------------------------------- Gilbert-Peierls LU (SparseLuFactorization.cs)
}  // namespace

struct orc_lu {
  int n = 0;
  vector<int> lp, li, up, ui, pinv, q;
  vector<double> lx, ux;
  // plan (keepPlan): reach per column in DFS topological order, pivot row per step, L rows as original rows
  vector<int> reach_ptr, reach, pivot_row, li_original;
  vector<int> plan_ap, plan_ai;
  double diagonal_preference = 0.0;
};

namespace {

// NumericTolerances.MaxAbs (:165) and RowPivotFloorsCsc (:99-115) with ToRowPivotFloors.
bool row_floors(const Csc& a, vector<double>& floors, double& anorm) {
  anorm = 0.0;
  int nnz = a.ap[a.n];
  for (int p = 0; p < nnz; p++) {
    double v = a.ax[p];
    if (std::isnan(v)) { anorm = NAN; return false; }
    double av = std::fabs(v);
    if (av > anorm) anorm = av;
  }
  if (std::isinf(anorm)) return false;
  floors.assign(a.n, 0.0);
  for (int p = 0; p < nnz; p++) {
    double av = std::fabs(a.ax[p]);
    if (av > floors[a.ai[p]]) floors[a.ai[p]] = av;
  }
  for (int r = 0; r < a.n; r++) floors[r] = floors[r] * kRelativePivotFloor;
  return true;
}

int dfs(int j, const vector<int>& lp, const vector<int>& li, const vector<int>& pinv, int top, vector<int>& xi,
        vector<int>& pstack, vector<char>& marked) {
  int head = 0;
  xi[0] = j;
  while (head >= 0) {
    j = xi[head];
    int jcol = pinv[j];
    if (!marked[j]) {
      marked[j] = 1;
      pstack[head] = jcol < 0 ? 0 : lp[jcol] + 1;
    }
    bool done = true;
    int pend = jcol < 0 ? 0 : lp[jcol + 1];
    for (int p = pstack[head]; p < pend; p++) {
      int i = li[p];
      if (marked[i]) continue;
      pstack[head] = p + 1;
      xi[++head] = i;
      done = false;
      break;
    }
    if (done) {
      head--;
      xi[--top] = j;
    }
  }
  return top;
}

int reach(int col, const Csc& a, const vector<int>& lp, const vector<int>& li, const vector<int>& pinv, int n,
          vector<int>& xi, vector<int>& pstack, vector<char>& marked) {
  int top = n;
  for (int p = a.ap[col]; p < a.ap[col + 1]; p++) {
    int j = a.ai[p];
    if (!marked[j]) top = dfs(j, lp, li, pinv, top, xi, pstack, marked);
  }
  for (int p = top; p < n; p++) marked[xi[p]] = 0;
  return top;
}

// Factor (SparseLuFactorization.cs:69-195) with keepPlan; returns null when singular. Non-finite entries are refused
// by the reference with an ArgumentException; here they return null with *nonfinite set.
orc_lu* gp_factor(const Csc& a, const int* q, double diag_pref, bool* nonfinite) {
  int n = a.n;
  vector<double> floors;
  double anorm;
  *nonfinite = false;
  if (!row_floors(a, floors, anorm)) { *nonfinite = true; return nullptr; }
  auto* f = new orc_lu();
  f->n = n;
  f->diagonal_preference = diag_pref;
  f->lp.assign(n + 1, 0);
  f->up.assign(n + 1, 0);
  f->pinv.assign(n, -1);
  f->q.assign(q, q + n);
  f->reach_ptr.assign(n + 1, 0);
  f->pivot_row.assign(n, -1);
  vector<double> x(n, 0.0);
  vector<int> xi(n), pstack(n);
  vector<char> marked(n, 0);
  vector<int>& li = f->li;
  vector<double>& lx = f->lx;
  for (int k = 0; k < n; k++) {
    f->lp[k] = (int)li.size();
    f->up[k] = (int)f->ui.size();
    int col = q[k];
    int top = reach(col, a, f->lp, li, f->pinv, n, xi, pstack, marked);
    // SpSolve: clear, scatter, update in topological order
    for (int p = top; p < n; p++) x[xi[p]] = 0.0;
    for (int p = a.ap[col]; p < a.ap[col + 1]; p++) x[a.ai[p]] = a.ax[p];
    for (int px = top; px < n; px++) {
      int j = xi[px];
      int jcol = f->pinv[j];
      if (jcol < 0) continue;
      double xj = x[j];
      for (int p = f->lp[jcol] + 1; p < f->lp[jcol + 1]; p++) x[li[p]] -= lx[p] * xj;
    }
    for (int p = top; p < n; p++) f->reach.push_back(xi[p]);
    f->reach_ptr[k + 1] = (int)f->reach.size();
    int ipiv = -1;
    double best = -1.0;
    for (int px = top; px < n; px++) {
      int i = xi[px];
      if (f->pinv[i] < 0) {
        double t = std::fabs(x[i]);
        if (t > best) { best = t; ipiv = i; }
      } else {
        f->ui.push_back(f->pinv[i]);
        f->ux.push_back(x[i]);
      }
    }
    if (ipiv < 0 || best <= floors[ipiv]) { delete f; return nullptr; }
    if (diag_pref > 0.0) {
      // PreferDiagonal (SparseLuFactorization.cs:201-213)
      if (!(ipiv == col || f->pinv[col] >= 0)) {
        for (int px = top; px < n; px++) {
          if (xi[px] != col) continue;
          double td = std::fabs(x[col]);
          if (td >= diag_pref * best && td > floors[col]) ipiv = col;
          break;
        }
      }
    }
    double pivot = x[ipiv];
    f->pivot_row[k] = ipiv;
    f->ui.push_back(k);
    f->ux.push_back(pivot);
    f->pinv[ipiv] = k;
    li.push_back(ipiv);
    lx.push_back(1.0);
    for (int px = top; px < n; px++) {
      int i = xi[px];
      if (f->pinv[i] < 0) {
        li.push_back(i);
        lx.push_back(x[i] / pivot);
      }
      x[i] = 0.0;
    }
  }
  f->lp[n] = (int)li.size();
  f->up[n] = (int)f->ui.size();
  f->li_original = li;
  for (size_t p = 0; p < li.size(); p++) li[p] = f->pinv[li[p]];
  f->plan_ap = a.ap;
  f->plan_ai = a.ai;
  return f;
}

// Refactor (SparseLuFactorization.cs:246-306): replays the plan; null when the pivot would move or is singular.
orc_lu* gp_refactor(const orc_lu& plan, const Csc& a) {
  int n = plan.n;
  vector<double> floors;
  double anorm;
  if (!row_floors(a, floors, anorm)) return nullptr;
  auto* f = new orc_lu(plan);
  f->lx.assign(plan.lp[n], 0.0);
  f->ux.assign(plan.up[n], 0.0);
  vector<double> x(n, 0.0);
  const vector<int>& step = plan.pinv;
  int lnz = 0, unz = 0;
  for (int k = 0; k < n; k++) {
    int r0 = plan.reach_ptr[k], r1 = plan.reach_ptr[k + 1];
    int col = plan.q[k];
    for (int p = r0; p < r1; p++) x[plan.reach[p]] = 0.0;
    for (int p = a.ap[col]; p < a.ap[col + 1]; p++) x[a.ai[p]] = a.ax[p];
    for (int px = r0; px < r1; px++) {
      int j = plan.reach[px];
      int jcol = step[j];
      if (jcol >= k) continue;
      double xj = x[j];
      for (int p = plan.lp[jcol] + 1; p < plan.lp[jcol + 1]; p++) x[plan.li_original[p]] -= f->lx[p] * xj;
    }
    int ipiv = -1;
    double best = -1.0;
    for (int px = r0; px < r1; px++) {
      int i = plan.reach[px];
      if (step[i] >= k) {
        double t = std::fabs(x[i]);
        if (t > best) { best = t; ipiv = i; }
      } else {
        f->ux[unz++] = x[i];
      }
    }
    bool clears = ipiv >= 0 && best > floors[ipiv];
    if (clears && plan.diagonal_preference > 0.0 && !(ipiv == col || step[col] < k)) {
      for (int px = r0; px < r1; px++) {
        if (plan.reach[px] != col) continue;
        double td = std::fabs(x[col]);
        if (td >= plan.diagonal_preference * best && td > floors[col]) ipiv = col;
        break;
      }
    }
    if (ipiv != plan.pivot_row[k] || !clears) { delete f; return nullptr; }
    double pivot = x[ipiv];
    f->ux[unz++] = pivot;
    f->lx[lnz++] = 1.0;
    for (int px = r0; px < r1; px++) {
      int i = plan.reach[px];
      if (step[i] > k) f->lx[lnz++] = x[i] / pivot;
      x[i] = 0.0;
    }
  }
  return f;
}

// Solve (SparseLuFactorization.cs:338-371).
void lu_solve(const orc_lu& f, const double* b, double* out) {
  int n = f.n;
  vector<double> x(n);
  for (int k = 0; k < n; k++) x[f.pinv[k]] = b[k];
  for (int j = 0; j < n; j++) {
    double xj = x[j];
    for (int p = f.lp[j] + 1; p < f.lp[j + 1]; p++) x[f.li[p]] -= f.lx[p] * xj;
  }
  for (int j = n - 1; j >= 0; j--) {
    double xj = x[j] / f.ux[f.up[j + 1] - 1];
    x[j] = xj;
    for (int p = f.up[j]; p < f.up[j + 1] - 1; p++) x[f.ui[p]] -= f.ux[p] * xj;
  }
  for (int k = 0; k < n; k++) out[f.q[k]] = x[k];
}

// ---------------------------------------------------------------- ILU(0) + BiCGSTAB (Sparse.cs:230-405)
struct Ilu0 {
  const Csr* a;
  vector<double> lu;
};

bool ilu0_factor(const Csr& a, Ilu0& out) {
  int n = a.n;
  for (int i = 0; i < n; i++) if (a.diag[i] < 0) return false;
  out.a = &a;
  out.lu = a.val;
  vector<int> iw(n, -1);
  for (int i = 0; i < n; i++) {
    for (int k = a.rp[i]; k < a.rp[i + 1]; k++) iw[a.ci[k]] = k;
    for (int k = a.rp[i]; k < a.diag[i]; k++) {
      int j = a.ci[k];
      double piv = out.lu[a.diag[j]];
      double lij = (piv != 0.0) ? out.lu[k] / piv : 0.0;
      out.lu[k] = lij;
      if (lij == 0.0) continue;
      for (int kk = a.diag[j] + 1; kk < a.rp[j + 1]; kk++) {
        int pos = iw[a.ci[kk]];
        if (pos != -1) out.lu[pos] -= lij * out.lu[kk];
      }
    }
    double row_scale = 0.0;
    for (int k = a.rp[i]; k < a.rp[i + 1]; k++) { double av = std::fabs(a.val[k]); if (av > row_scale) row_scale = av; }
    double guard = kRelativePivotFloor * (row_scale > 0.0 ? row_scale : 1.0);
    double d = out.lu[a.diag[i]];
    if (std::fabs(d) < guard) out.lu[a.diag[i]] = d < 0 ? -guard : guard;
    for (int k = a.rp[i]; k < a.rp[i + 1]; k++) iw[a.ci[k]] = -1;
  }
  return true;
}

void ilu0_apply(const Ilu0& m, const double* r, double* z) {
  const Csr& a = *m.a;
  for (int i = 0; i < a.n; i++) {
    double s = r[i];
    for (int k = a.rp[i]; k < a.diag[i]; k++) s -= m.lu[k] * z[a.ci[k]];
    z[i] = s;
  }
  for (int i = a.n - 1; i >= 0; i--) {
    double s = z[i];
    for (int k = a.diag[i] + 1; k < a.rp[i + 1]; k++) s -= m.lu[k] * z[a.ci[k]];
    z[i] = s / m.lu[a.diag[i]];
  }
}

double dot(const vector<double>& a, const vector<double>& b) {
  double s = 0;
  for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
  return s;
}

bool bicgstab(const Csr& a, const vector<double>& b, vector<double>& x, const Ilu0& pc, double tol, int max_iter) {
  int n = a.n;
  std::fill(x.begin(), x.end(), 0.0);
  vector<double> r(b), rhat(b), p(n, 0), v(n, 0), s(n, 0), t(n, 0), ph(n, 0), sh(n, 0);
  double bnorm = std::sqrt(dot(b, b));
  if (bnorm == 0) {
    for (double bi : b) if (bi != 0.0) return false;
    return true;
  }
  auto true_residual = [&]() {
    csr_multiply(a, x.data(), t.data());
    double sum = 0.0;
    for (int i = 0; i < n; i++) { double d = b[i] - t[i]; sum += d * d; }
    return std::sqrt(sum) / bnorm < tol;
  };
  double rho = 1, alpha = 1, omega = 1;
  for (int it = 0; it < max_iter; it++) {
    double rho_new = dot(rhat, r);
    if (std::fabs(rho_new) < kUnderflowGuard || std::isnan(rho_new)) return false;
    double beta = (rho_new / rho) * (alpha / omega);
    for (int i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    ilu0_apply(pc, p.data(), ph.data());
    csr_multiply(a, ph.data(), v.data());
    double rhat_v = dot(rhat, v);
    if (rhat_v == 0.0 || std::isnan(rhat_v)) return false;
    alpha = rho_new / rhat_v;
    for (int i = 0; i < n; i++) s[i] = r[i] - alpha * v[i];
    if (std::sqrt(dot(s, s)) / bnorm < tol) {
      for (int i = 0; i < n; i++) x[i] += alpha * ph[i];
      return true_residual();
    }
    ilu0_apply(pc, s.data(), sh.data());
    csr_multiply(a, sh.data(), t.data());
    double tt = dot(t, t);
    omega = tt > 0 ? dot(t, s) / tt : 0.0;
    for (int i = 0; i < n; i++) { x[i] += alpha * ph[i] + omega * sh[i]; r[i] = s[i] - omega * t[i]; }
    rho = rho_new;
    if (std::sqrt(dot(r, r)) / bnorm < tol) return true_residual();
    if (omega == 0) return false;
  }
  return false;
}

// ---------------------------------------------------------------- CpuSparseLuSolver (ordering cached per pattern)
struct CpuSparseLuSolver {
  vector<int> q, rp, ci;
  bool have = false;
  // 0 ok, 1 singular (null), 2 non-finite matrix or right-hand side (ArgumentException in the reference)
  int solve(const Csr& a, const double* b, double* x) {
    if (a.n == 0) return 0;
    if (!have || rp != a.rp || ci != a.ci) {
      q = md_order(a.n, a.rp.data(), a.ci.data());
      rp = a.rp;
      ci = a.ci;
      have = true;
    }
    Csc c = to_csc(a.n, a.rp.data(), a.ci.data(), a.val.data(), false);
    bool nonfinite = false;
    orc_lu* f = gp_factor(c, q.data(), 0.0, &nonfinite);
    if (nonfinite) return 2;
    if (!f) return 1;
    for (int k = 0; k < a.n; k++)
      if (!std::isfinite(b[k])) { delete f; return 2; }
    lu_solve(*f, b, x);
    delete f;
    return 0;
  }
};

// ---------------------------------------------------------------- the Newton loop (RunNewton, Q limits off)
struct NewtonOut {
  bool converged = false, diverged = false;
  int iters = 0;
  double max_mis = INFINITY;
  int step_failures = 0;
  double start_mis = NAN, last_from = NAN;   // NewtonShell.NewtonRunEnd's StartMismatch, BeforeLastStepMismatch
  bool cap_reached = false;
};

NewtonOut run_newton(const Ycsr& y, const vector<int>& type, vector<double>& vm, vector<double>& va,
                     const vector<double>& pspec, const vector<double>& qspec, vector<double>& p, vector<double>& q,
                     const orc_options& opt) {
  int n = y.n;
  vector<int> ppos, qpos;
  int m = equation_positions(type, ppos, qpos);
  vector<double> f(m), dx(m);
  vector<int> slots;
  Csr jac = build_jacobian_structure(y, ppos, qpos, m, slots);
  CpuSparseLuSolver solver;
  NewtonOut out;
  double start_mis = NAN;
  int iter = 0;
  for (;; iter++) {
    compute_injections(y, vm.data(), va.data(), p.data(), q.data());
    for (int i = 0; i < n; i++) {
      if (ppos[i] >= 0) f[ppos[i]] = pspec[i] - p[i];
      if (qpos[i] >= 0) f[qpos[i]] = qspec[i] - q[i];
    }
    out.max_mis = 0;
    for (int k = 0; k < m; k++) out.max_mis = net_max(out.max_mis, std::fabs(f[k]));
    if (iter == 0) start_mis = out.max_mis;
    if (out.max_mis < opt.tolerance) { out.converged = true; break; }
    if (out.max_mis > net_max(opt.diverge_limit, start_mis)) { out.diverged = true; break; }
    if (iter >= opt.max_iterations) break;
    refill_jacobian(jac, slots, y, vm.data(), va.data(), p.data(), q.data());
    // NewtonShell.SolveStep
    bool ok = true;
    for (double jv : jac.val) if (!std::isfinite(jv)) { ok = false; break; }
    if (ok) for (double fv : f) if (!std::isfinite(fv)) { ok = false; break; }
    if (!ok) break;
    int rc = solver.solve(jac, f.data(), dx.data());
    if (rc == 2) break;
    if (rc == 1) {
      out.step_failures++;
      Ilu0 pc;
      vector<double> xs(m, 0.0), fv(f);
      if (!ilu0_factor(jac, pc)) break;
      double tol = std::min(1e-10, opt.tolerance * 1e-3);
      if (!bicgstab(jac, fv, xs, pc, tol, 4 * m + 200)) break;
      dx = xs;
    }
    double max_abs = 0;
    bool finite = true;
    for (int k = 0; k < m; k++) {
      if (!std::isfinite(dx[k])) { finite = false; break; }
      double av = std::fabs(dx[k]);
      if (av > max_abs) max_abs = av;
    }
    if (!finite) break;
    out.last_from = out.max_mis;   // the retry's stall test only (NewtonShell.Stalled)
    // NewtonShell.ApplyStep
    double damp = max_abs > kMaxUndampedStep ? kMaxUndampedStep / max_abs : 1.0;
    for (int i = 0; i < n; i++) {
      if (ppos[i] >= 0) va[i] += damp * dx[ppos[i]];
      if (qpos[i] >= 0) { vm[i] += damp * dx[qpos[i]]; if (vm[i] < kVmFloorPu) vm[i] = kVmFloorPu; }
    }
  }
  out.iters = iter;
  out.start_mis = start_mis;
  out.cap_reached = !out.converged && !out.diverged && iter >= opt.max_iterations;
  return out;
}

// NewtonShell.ApplyQLimits (NewtonShell.cs:519-558) with RevertQLimits off, then RevertBrokenPins (:583-595): a bus
// that started PV and is still PV switches to PQ at the limit its generators' reactive output crossed (by more than
// QLimitTolMva); once nothing more switches, a pinned bus whose voltage lies on the wrong side of its set point (by
// more than RevertVoltTolPu) goes back to PV at its set point, at most once per solve. Returns whether any bus changed.
enum { kPinNone = 0, kPinMax = 1, kPinMin = 2, kPinFixed = 3 };
constexpr double kQLimitTolMva = 1e-3, kRevertVoltTolPu = 1e-4;

bool apply_q_limits(const BusSpec& s, vector<int>& eff, const vector<double>& q, vector<double>& qspec,
                    vector<double>& vm, vector<int>& pinned, vector<int>& reverts, double base) {
  const int n = (int)eff.size();
  bool changed = false;
  for (int i = 0; i < n; i++) {
    if (s.type0[i] != ORC_PV || eff[i] != ORC_PV) continue;
    const double qgen = q[i] * base + s.qd[i];
    if (qgen > s.qmax[i] + kQLimitTolMva) {
      eff[i] = ORC_PQ;
      qspec[i] = (s.qmax[i] - s.qd[i]) / base;
      pinned[i] = s.qmin[i] == s.qmax[i] ? kPinFixed : kPinMax;
      changed = true;
    } else if (qgen < s.qmin[i] - kQLimitTolMva) {
      eff[i] = ORC_PQ;
      qspec[i] = (s.qmin[i] - s.qd[i]) / base;
      pinned[i] = s.qmin[i] == s.qmax[i] ? kPinFixed : kPinMin;
      changed = true;
    }
  }
  if (changed) return true;
  for (int i = 0; i < n; i++) {
    if (s.type0[i] != ORC_PV || eff[i] != ORC_PQ) continue;
    if (pinned[i] == kPinNone || pinned[i] == kPinFixed || reverts[i] >= 1) continue;
    const bool rmax = pinned[i] == kPinMax && vm[i] > s.vg[i] + kRevertVoltTolPu;
    const bool rmin = pinned[i] == kPinMin && vm[i] < s.vg[i] - kRevertVoltTolPu;
    if (!(rmax || rmin)) continue;
    eff[i] = ORC_PV;
    vm[i] = s.vg[i];
    pinned[i] = kPinNone;
    reverts[i]++;
    changed = true;
  }
  return changed;
}

// ---------------------------------------------------------------- the DC-angle start retry (NewtonShell.cs:142-299)
constexpr double kRetryStallFactor = 1e-3, kRetryLastStepFactor = 0.5, kMaxRetryBranchAngleDeg = 90.0;

// NewtonShell.Stalled(NewtonRunEnd, double) (:200-206) over Stalled(double, double) (:162).
bool stalled(const NewtonOut& run, double final_mis) {
  if (run.cap_reached && run.iters == 0) return false;
  if (final_mis <= kRetryStallFactor * run.start_mis) return false;
  if (run.cap_reached && run.iters >= 2 && final_mis <= kRetryLastStepFactor * run.last_from) return false;
  return true;
}

// NewtonShell.TrySeedDcAngles (:282-299) with DcPowerFlow.Solve(DcDemand.LegacyView(model)) (DcPowerFlow.cs:78-86,
// DcDemand.cs:38-50) on DistributionFactors.Prepare / IslandsCore / DcSolve (DistributionFactors.cs:143-360,454-520):
// the reactance-only B' (every tap 1, every shift 0, every Gs 0: b = 1 / x per in-service branch, stamped in branch
// order with the reference bus's row and column left out), factored by SparseLu.Factor (minimum degree, explicit zeros
// dropped, partial pivoting); the right-hand side is each bus's net injection (0 - Pd, then + Pg per generator in
// generator order) / base; the angles are reduced to [-180, 180] degrees (IEEE remainder) and offset by the slack's
// start angle. False, spec unchanged, unless the case has exactly one slack bus, its DC model forms a single island
// (a connected graph over the in-service branches, none of zero reactance) and B' factors.
bool try_seed_dc_angles(const orc_network& net, BusSpec& spec) {
  const int n = net.n_bus;
  int slack = -1, slacks = 0;
  for (int i = 0; i < n; i++)
    if (spec.type0[i] == ORC_SLACK) { slack = i; slacks++; }
  if (slacks != 1) return false;
  // islands: union-find over the branches the DC model stamps; an in-service zero-reactance branch is refused
  vector<int> parent(n);
  for (int i = 0; i < n; i++) parent[i] = i;
  auto find = [&](int x) { while (parent[x] != x) { parent[x] = parent[parent[x]]; x = parent[x]; } return x; };
  for (int l = 0; l < net.n_branch; l++) {
    if (!net.br_in_service[l]) continue;
    if (net.br_x[l] == 0.0) return false;
    const int rf = find(net.br_from[l]), rt = find(net.br_to[l]);
    if (rf != rt) parent[rf] = rt;
  }
  for (int i = 0; i < n; i++)
    if (find(i) != find(slack)) return false;   // more than one island
  const int nr = n - 1;
  if (nr == 0) return false;                     // SparseLu.Factor of an empty B' is null
  vector<int> red(n);
  for (int i = 0, r = 0; i < n; i++) red[i] = i == slack ? -1 : r++;
  CsrBuilder bld(nr);
  for (int l = 0; l < net.n_branch; l++) {
    if (!net.br_in_service[l]) continue;
    const int f = net.br_from[l], t = net.br_to[l];
    const double bl = 1.0 / net.br_x[l];
    const bool ff = red[f] >= 0, tf = red[t] >= 0;
    if (ff) bld.add(red[f], red[f], bl);
    if (tf) bld.add(red[t], red[t], bl);
    if (ff && tf) { bld.add(red[f], red[t], -bl); bld.add(red[t], red[f], -bl); }
  }
  Csr bp = bld.build();
  vector<int> q = md_order(nr, bp.rp.data(), bp.ci.data());
  Csc c = to_csc(nr, bp.rp.data(), bp.ci.data(), bp.val.data(), false);
  bool nonfinite = false;
  orc_lu* f = gp_factor(c, q.data(), 0.0, &nonfinite);
  if (!f) return false;
  // DcPowerFlow.NetInjections: per bus 0 - Pd, then + Pg per generator in generator order (MW)
  vector<double> inj(n);
  for (int i = 0; i < n; i++) inj[i] = 0.0 - net.pd[i];
  for (int g = 0; g < net.n_gen; g++) inj[net.gen_bus[g]] += net.gen_pg[g];
  vector<double> rhs(nr, 0.0), x(nr);
  for (int i = 0; i < n; i++)
    if (red[i] >= 0) rhs[red[i]] += inj[i] / net.base_mva;
  lu_solve(*f, rhs.data(), x.data());
  delete f;
  const double va0 = spec.va[slack];
  for (int i = 0; i < n; i++) {
    if (i == slack) continue;
    const double va_deg = x[red[i]] * 180.0 / kPi;
    spec.va[i] = va0 + std::remainder(va_deg, 360.0) * kPi / 180.0;
  }
  return true;
}

// NewtonShell.MaxBranchAngleDeg (:311-320): the largest |theta_from - shift - theta_to| (degrees, wrapped) over the
// in-service branches.
double max_branch_angle_deg(const orc_network& net, const vector<double>& va) {
  double worst = 0;
  for (int l = 0; l < net.n_branch; l++) {
    if (!net.br_in_service[l]) continue;
    const double df = va[net.br_from[l]] * 180.0 / kPi, dt = va[net.br_to[l]] * 180.0 / kPi;
    const double d = std::fabs(std::remainder(df - net.br_shift_deg[l] - dt, 360.0));
    if (d > worst) worst = d;
  }
  return worst;
}

struct Attempt {
  orc_result res{};
  vector<double> vm, va;
  bool newton_failed = false;
  NewtonOut end;   // how the last Newton run ended
};

// SparseNewtonRaphsonPowerFlow.SolveAcAttempt (:71-165): one complete solve (reactive-limit passes included), from the
// ordinary start or, with dc_start, from the DC angles; false when that start cannot be formed.
bool solve_attempt(const orc_network& net, int skip, const orc_options& opt, bool dc_start, Attempt& a) {
  Ycsr y = build_y(net, skip);
  BusSpec spec = build_bus_spec(net, opt.flat_start != 0);
  if (dc_start && !try_seed_dc_angles(net, spec)) return false;
  int n = net.n_bus;
  vector<double> p(n), q(n);
  vector<int> eff(spec.type0), pinned(n, kPinNone), reverts(n, 0);
  const bool enforce = opt.enforce_q_limits != 0;
  const int passes = enforce ? std::max(1, opt.max_q_passes) : 1;
  int total = 0, failures = 0;
  bool pending = false;
  NewtonOut o;
  for (int qp = 0; qp < passes; qp++) {
    pending = false;
    o = run_newton(y, eff, spec.vm, spec.va, spec.pspec, spec.qspec, p, q, opt);
    total += o.iters;
    failures += o.step_failures;
    a.newton_failed = !o.converged;
    if (!enforce || !o.converged) break;
    pending = apply_q_limits(spec, eff, q, spec.qspec, spec.vm, pinned, reverts, net.base_mva);
    if (!pending) break;
  }
  if (pending) {
    // the cap came right after a switch: solve the switched types once more and report the state not converged
    o = run_newton(y, eff, spec.vm, spec.va, spec.pspec, spec.qspec, p, q, opt);
    total += o.iters;
    failures += o.step_failures;
    a.newton_failed = !o.converged;
    o.converged = false;
  }
  a.end = o;
  compute_injections(y, spec.vm.data(), spec.va.data(), p.data(), q.data());
  bool finite = std::isfinite(o.max_mis);
  for (int i = 0; i < n && finite; i++)
    finite = std::isfinite(spec.vm[i]) && std::isfinite(spec.va[i]) && std::isfinite(p[i] * net.base_mva) &&
             std::isfinite(q[i] * net.base_mva);
  // AssembleResult: diverged only for a run the divergence stop ended
  a.res.status = !finite ? ORC_NONFINITE : o.converged ? ORC_CONVERGED : o.diverged ? ORC_DIVERGED
                                                                                     : ORC_MAX_ITERATIONS;
  a.res.iterations = total;
  a.res.max_mismatch = o.max_mis;
  a.res.step_failures = failures;
  a.res.dc_retry = ORC_RETRY_NOT_TRIED;
  a.vm = std::move(spec.vm);
  a.va = std::move(spec.va);
  return true;
}

// NewtonShell.SolveWithRetry (:227-268): the ordinary solve, then (only when its last Newton run did not converge,
// stalled or diverged, and the options allow it) the solve from the DC-angle start, whose state is reported when it
// converged with every branch angle within 90 degrees.
void solve_one(const orc_network& net, int skip, const orc_options& opt, double* vm_out, double* va_out,
               orc_result* res) {
  auto t0 = std::chrono::steady_clock::now();
  Attempt first;
  solve_attempt(net, skip, opt, false, first);
  const Attempt* out = &first;
  Attempt retry;
  if (first.res.status != ORC_CONVERGED && first.newton_failed && opt.retry_dc_start && opt.max_iterations > 1 &&
      stalled(first.end, first.res.max_mismatch)) {
    if (!solve_attempt(net, skip, opt, true, retry)) {
      first.res.dc_retry = ORC_RETRY_UNAVAILABLE;
    } else if (retry.res.status == ORC_CONVERGED) {
      if (max_branch_angle_deg(net, retry.va) <= kMaxRetryBranchAngleDeg) {
        retry.res.dc_retry = ORC_RETRY_CONVERGED;
        out = &retry;
      } else {
        first.res.dc_retry = ORC_RETRY_REJECTED;
      }
    } else {
      first.res.dc_retry = ORC_RETRY_DID_NOT_CONVERGE;
    }
  }
  *res = out->res;
  const int n = net.n_bus;
  if (vm_out) std::memcpy(vm_out, out->vm.data(), sizeof(double) * n);
  if (va_out) std::memcpy(va_out, out->va.data(), sizeof(double) * n);
  res->solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// A copy of the network with the given branches out of service: the oracle's WithoutMany (ContingencyAnalysis.cs:
// 1839-1845; the Y-bus builder, the line flows and the DC model all skip an out-of-service branch, so removing a branch
// and switching it out give the same numbers). `mask` holds the copy's in-service flags.
orc_network without(const orc_network& net, const int32_t* skips, int n_skip, vector<uint8_t>& mask) {
  mask.assign(net.br_in_service, net.br_in_service + net.n_branch);
  for (int k = 0; k < n_skip; k++)
    if (skips[k] >= 0 && skips[k] < net.n_branch) mask[skips[k]] = 0;
  orc_network c = net;
  c.br_in_service = mask.data();
  return c;
}

}  // namespace

// ================================================================ C ABI
extern "C" {

void orc_sincos(const double* x, int64_t n, double* s, double* c) {
  for (int64_t i = 0; i < n; i++) orc_fdlibm::sincos(x[i], &s[i], &c[i]);
}

void orc_complex_div(double a, double b, double c, double d, double* re, double* im) {
  Cx z = cx_div(Cx{a, b}, Cx{c, d});
  *re = z.re;
  *im = z.im;
}

void orc_branch_stamp(double r, double x, double b, double tap, double shift_deg, double* out8) {
  Stamp s = branch_stamp(r, x, b, tap, shift_deg);
  Cx v[4] = {s.yff, s.ytt, s.yft, s.ytf};
  for (int i = 0; i < 4; i++) { out8[2 * i] = v[i].re; out8[2 * i + 1] = v[i].im; }
}

int32_t orc_ybus(const orc_network* net, int32_t skip, int32_t* rowptr, int32_t* colidx, double* g, double* b) {
  Ycsr y = build_y(*net, skip);
  if (rowptr) std::memcpy(rowptr, y.rp.data(), sizeof(int32_t) * (y.n + 1));
  if (colidx) std::memcpy(colidx, y.ci.data(), sizeof(int32_t) * y.ci.size());
  if (g) std::memcpy(g, y.g.data(), sizeof(double) * y.g.size());
  if (b) std::memcpy(b, y.b.data(), sizeof(double) * y.b.size());
  return (int32_t)y.ci.size();
}

void orc_bus_spec(const orc_network* net, int32_t flat_start, int32_t* type, double* vm, double* va, double* pspec,
                  double* qspec) {
  BusSpec s = build_bus_spec(*net, flat_start != 0);
  for (int i = 0; i < net->n_bus; i++) {
    type[i] = s.type0[i]; vm[i] = s.vm[i]; va[i] = s.va[i]; pspec[i] = s.pspec[i]; qspec[i] = s.qspec[i];
  }
}

int32_t orc_mismatch(const orc_network* net, int32_t skip, const double* vm, const double* va, double* p, double* q,
                     double* f) {
  Ycsr y = build_y(*net, skip);
  BusSpec s = build_bus_spec(*net, false);
  vector<int> ppos, qpos;
  int m = equation_positions(s.type0, ppos, qpos);
  compute_injections(y, vm, va, p, q);
  for (int i = 0; i < net->n_bus; i++) {
    if (ppos[i] >= 0) f[ppos[i]] = s.pspec[i] - p[i];
    if (qpos[i] >= 0) f[qpos[i]] = s.qspec[i] - q[i];
  }
  return m;
}

int32_t orc_jacobian(const orc_network* net, int32_t skip, const double* vm, const double* va, int32_t* m_out,
                     int32_t* rowptr, int32_t* colidx, double* val) {
  Ycsr y = build_y(*net, skip);
  BusSpec s = build_bus_spec(*net, false);
  vector<int> ppos, qpos, slots;
  int m = equation_positions(s.type0, ppos, qpos);
  Csr jac = build_jacobian_structure(y, ppos, qpos, m, slots);
  if (m_out) *m_out = m;
  if (val) {
    vector<double> p(net->n_bus), q(net->n_bus);
    compute_injections(y, vm, va, p.data(), q.data());
    refill_jacobian(jac, slots, y, vm, va, p.data(), q.data());
    std::memcpy(val, jac.val.data(), sizeof(double) * jac.val.size());
  }
  if (rowptr) std::memcpy(rowptr, jac.rp.data(), sizeof(int32_t) * (m + 1));
  if (colidx) std::memcpy(colidx, jac.ci.data(), sizeof(int32_t) * jac.ci.size());
  return (int32_t)jac.ci.size();
}

// NewtonShell.ComputeLineFlows (NewtonShell.cs:699-734): per branch the from / to end complex powers in MVA and the
// loading against rate A (percent, 0 when unrated). Index-aligned with the model's branches; an out-of-service branch
// and the outaged one (skip) carry no flow. pf, qf, pt, qt, loading: [n_branch] each (any may be null).
void orc_line_flows(const orc_network* net, int32_t skip, const double* rate_a, const double* vm, const double* va,
                    double* pf, double* qf, double* pt, double* qt, double* loading) {
  for (int k = 0; k < net->n_branch; k++) {
    double v[5] = {0, 0, 0, 0, 0};
    if (net->br_in_service[k] && k != skip) {
      const int f = net->br_from[k], t = net->br_to[k];
      double sf, cf, st, ct;
      orc_fdlibm::sincos(va[f], &sf, &cf);
      orc_fdlibm::sincos(va[t], &st, &ct);
      Cx vf = {vm[f] * cf, vm[f] * sf}, vt = {vm[t] * ct, vm[t] * st};  // DetMath.FromPolar
      Stamp s = branch_stamp(net->br_r[k], net->br_x[k], net->br_b[k], net->br_tap[k], net->br_shift_deg[k]);
      Cx i_f = cx_add(cx_mul(s.yff, vf), cx_mul(s.yft, vt));
      Cx i_t = cx_add(cx_mul(s.ytf, vf), cx_mul(s.ytt, vt));
      Cx s_f = cx_scale(cx_mul(vf, cx_conj(i_f)), net->base_mva);
      Cx s_t = cx_scale(cx_mul(vt, cx_conj(i_t)), net->base_mva);
      const double mva = net_max(cx_abs(s_f), cx_abs(s_t));
      const double rate = rate_a ? rate_a[k] : 0.0;
      v[0] = s_f.re; v[1] = s_f.im; v[2] = s_t.re; v[3] = s_t.im;
      v[4] = rate > 0 ? mva / rate * 100.0 : 0.0;
    }
    if (pf) pf[k] = v[0];
    if (qf) qf[k] = v[1];
    if (pt) pt[k] = v[2];
    if (qt) qt[k] = v[3];
    if (loading) loading[k] = v[4];
  }
}

// The numbers ContingencyAnalysis.ScreenStateCore classifies a converged outage by (ContingencyAnalysis.cs:1150-1236)
// and AssembleResult's losses (NewtonShell.cs:649-690): out[0] max branch loading (percent), out[1] branches above
// 100 %, out[2] min Vm, out[3] max Vm, out[4] buses below their low limit, out[5] buses above their high limit,
// out[6] total loss MW (sum of the bus P injections), out[7] total loss MVAr. p, q: the per-unit injections at the
// solved state. bus_low / bus_high: each bus's own limits (VoltageLimits.OperationalLowFor / HighFor: NaN = none, the
// operational band 0.94 / 1.06 pu applies on that side), or null for the band everywhere.
void orc_screen_summary(const orc_network* net, int32_t skip, const double* rate_a, const double* bus_low,
                        const double* bus_high, const double* vm, const double* va, const double* p, const double* q,
                        double* out) {
  vector<double> loading(net->n_branch);
  orc_line_flows(net, skip, rate_a, vm, va, nullptr, nullptr, nullptr, nullptr, loading.data());
  double max_load = 0, over = 0;
  for (int k = 0; k < net->n_branch; k++) {
    if (loading[k] > max_load) max_load = loading[k];
    if (loading[k] > 100.0) over += 1;
  }
  double vmin = vm[0], vmax = vm[0], under_n = 0, over_n = 0, lp = 0, lq = 0;
  for (int i = 0; i < net->n_bus; i++) {
    vmin = vm[i] < vmin ? vm[i] : vmin;
    vmax = vm[i] > vmax ? vm[i] : vmax;
    const double lo = bus_low && !std::isnan(bus_low[i]) ? bus_low[i] : 0.94;
    const double hi = bus_high && !std::isnan(bus_high[i]) ? bus_high[i] : 1.06;
    if (vm[i] < lo) under_n += 1;
    else if (vm[i] > hi) over_n += 1;
    lp += p[i] * net->base_mva;
    lq += q[i] * net->base_mva;
  }
  out[0] = max_load; out[1] = over; out[2] = vmin; out[3] = vmax; out[4] = under_n; out[5] = over_n;
  out[6] = lp; out[7] = lq;
}

void orc_md_order(int32_t n, const int32_t* rowptr, const int32_t* colidx, int32_t* q) {
  vector<int> o = md_order(n, rowptr, colidx);
  std::memcpy(q, o.data(), sizeof(int32_t) * n);
}

orc_lu* orc_lu_factor(int32_t n, const int32_t* rowptr, const int32_t* colidx, const double* val, const int32_t* q,
                      int32_t keep_zeros, double diagonal_preference) {
  Csc c = to_csc(n, rowptr, colidx, val, keep_zeros != 0);
  bool nonfinite = false;
  return gp_factor(c, q, diagonal_preference, &nonfinite);
}

void orc_lu_sizes(const orc_lu* h, int32_t* lnz, int32_t* unz, int32_t* reach_nnz) {
  *lnz = h->lp[h->n];
  *unz = h->up[h->n];
  *reach_nnz = (int32_t)h->reach.size();
}

void orc_lu_get(const orc_lu* h, int32_t* lp, int32_t* li, double* lx, int32_t* up, int32_t* ui, double* ux,
                int32_t* pinv, int32_t* q, int32_t* pivot_row, int32_t* reach_ptr, int32_t* reach) {
  int n = h->n;
  auto cp = [](auto* dst, const auto& src) { if (dst) std::memcpy(dst, src.data(), sizeof(src[0]) * src.size()); };
  cp(lp, h->lp); cp(li, h->li); cp(lx, h->lx); cp(up, h->up); cp(ui, h->ui); cp(ux, h->ux);
  cp(pinv, h->pinv); cp(q, h->q); cp(pivot_row, h->pivot_row); cp(reach_ptr, h->reach_ptr); cp(reach, h->reach);
  (void)n;
}

void orc_lu_solve(const orc_lu* h, const double* b, double* x) { lu_solve(*h, b, x); }

orc_lu* orc_lu_refactor(const orc_lu* plan, int32_t n, const int32_t* rowptr, const int32_t* colidx,
                        const double* val) {
  Csc c = to_csc(n, rowptr, colidx, val, true);
  if (c.ap != plan->plan_ap || c.ai != plan->plan_ai) return nullptr;  // SparseLuPlan.SamePattern
  return gp_refactor(*plan, c);
}

void orc_lu_free(orc_lu* h) { delete h; }

int32_t orc_sparse_solve(int32_t n, const int32_t* rowptr, const int32_t* colidx, const double* val,
                         const double* b, double* x) {
  Csr a;
  a.n = n;
  a.rp.assign(rowptr, rowptr + n + 1);
  a.ci.assign(colidx, colidx + rowptr[n]);
  a.val.assign(val, val + rowptr[n]);
  CpuSparseLuSolver s;
  return s.solve(a, b, x);
}

static Csr csr_with_diag(int32_t n, const int32_t* rowptr, const int32_t* colidx, const double* val) {
  Csr a;
  a.n = n;
  a.rp.assign(rowptr, rowptr + n + 1);
  a.ci.assign(colidx, colidx + rowptr[n]);
  a.val.assign(val, val + rowptr[n]);
  a.diag.assign(n, -1);
  for (int i = 0; i < n; i++)
    for (int k = a.rp[i]; k < a.rp[i + 1]; k++)
      if (a.ci[k] == i) a.diag[i] = k;
  return a;
}

int32_t orc_ilu0(int32_t n, const int32_t* rowptr, const int32_t* colidx, const double* val, double* lu_out) {
  Csr a = csr_with_diag(n, rowptr, colidx, val);
  Ilu0 pc;
  if (!ilu0_factor(a, pc)) return 0;
  std::copy(pc.lu.begin(), pc.lu.end(), lu_out);
  return 1;
}

int32_t orc_bicgstab(int32_t n, const int32_t* rowptr, const int32_t* colidx, const double* val, const double* b,
                     double tol, int32_t max_iter, double* x) {
  Csr a = csr_with_diag(n, rowptr, colidx, val);
  Ilu0 pc;
  if (!ilu0_factor(a, pc)) return 0;
  vector<double> bv(b, b + n), xv(n, 0.0);
  const bool ok = bicgstab(a, bv, xv, pc, tol, max_iter);
  std::copy(xv.begin(), xv.end(), x);
  return ok ? 1 : 0;
}

void orc_solve(const orc_network* net, int32_t skip, const orc_options* opt, double* vm_out, double* va_out,
               orc_result* res) {
  solve_one(*net, skip, *opt, vm_out, va_out, res);
}

void orc_screen_sets(const orc_network* net, const int32_t* offsets, const int32_t* branches, int32_t count,
                     const orc_options* opt, int32_t threads, orc_result* res, double* vm_out, double* va_out) {
  if (threads < 1) threads = 1;
  std::atomic<int> next{0};
  auto worker = [&]() {
    vector<uint8_t> mask;
    for (;;) {
      int k = next.fetch_add(1);
      if (k >= count) break;
      size_t off = (size_t)k * net->n_bus;
      orc_network c = without(*net, branches + offsets[k], offsets[k + 1] - offsets[k], mask);
      solve_one(c, -1, *opt, vm_out ? vm_out + off : nullptr, va_out ? va_out + off : nullptr, &res[k]);
    }
  };
  vector<std::thread> pool;
  for (int t = 0; t < threads; t++) pool.emplace_back(worker);
  for (auto& t : pool) t.join();
}

int32_t orc_dc_seed(const orc_network* net, int32_t flat_start, double* va) {
  BusSpec spec = build_bus_spec(*net, flat_start != 0);
  if (!try_seed_dc_angles(*net, spec)) return 0;
  std::memcpy(va, spec.va.data(), sizeof(double) * net->n_bus);
  return 1;
}

void orc_screen(const orc_network* net, const int32_t* outages, int32_t count, const orc_options* opt,
                int32_t threads, orc_result* res, double* vm_out, double* va_out) {
  if (threads < 1) threads = 1;
  std::atomic<int> next{0};
  auto worker = [&]() {
    for (;;) {
      int k = next.fetch_add(1);
      if (k >= count) break;
      size_t off = (size_t)k * net->n_bus;
      solve_one(*net, outages[k], *opt, vm_out ? vm_out + off : nullptr, va_out ? va_out + off : nullptr, &res[k]);
    }
  };
  vector<std::thread> pool;
  for (int t = 0; t < threads; t++) pool.emplace_back(worker);
  for (auto& t : pool) t.join();
}

}  // extern "C"
