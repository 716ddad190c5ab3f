// Host-side model and symbolic analysis (see model.h). Reference files restated:
//   AdmittanceStamp.cs:35-75 (stamp), SparseNewtonRaphsonPowerFlow.cs:382-402 (BuildY), Sparse.cs:145-222
//   (ComplexCsr.Builder: stable (column, insertion) order, duplicates summed from +0.0, zero diagonal injected),
//   NewtonShell.cs:433-486 (BuildBusSpec), :327-336 (EquationPositions), SparseNewtonRaphsonPowerFlow.cs:269-380
//   (injections, Jacobian structure and refill), MinimumDegreeOrdering.cs:23-136, SparseLuFactorization.cs:69-195
//   and :753-821 (Gilbert-Peierls with the plan), NumericTolerances.cs:36,99-115 (row pivot floors).
#include "model.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>

#include "fdlibm.h"

namespace gs {

namespace {
constexpr double kPi = 3.141592653589793;

inline Cx add(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
// System.Numerics.Complex division (Smith's formula, as the .NET BCL operator computes it).
inline Cx cdiv(Cx l, Cx r) {
  double a = l.re, b = l.im, c = r.re, d = r.im;
  if (std::fabs(d) < std::fabs(c)) {
    double doc = d / c;
    return {(a + b * doc) / (c + d * doc), (b - a * doc) / (c + d * doc)};
  }
  double cod = c / d;
  return {(b + a * cod) / (d + c * cod), (-a + b * cod) / (d + c * cod)};
}
inline Cx cdiv_real(Cx l, double r) {
  if (r == 0) return {std::nan(""), std::nan("")};
  return {l.re / r, l.im / r};
}
}  // namespace

Stamp branch_stamp(double r, double x, double b, double tap_ratio, double shift_deg) {
  Cx ys = cdiv(Cx{1.0, 0.0}, Cx{r * 1.0, x * 1.0});
  Cx bc{0.0, (b * 1.0) / 2.0};
  double a = tap_ratio == 0 ? 1.0 : tap_ratio;
  double shift = shift_deg * kPi / 180.0;
  double s, c;
  gs_fdlibm::sincos(shift, &s, &c);
  Cx tap{a * c, a * s};
  Stamp st;
  st.yff = cdiv_real(add(ys, bc), a * a);
  st.ytt = add(ys, bc);
  st.yft = cdiv(Cx{-ys.re, -ys.im}, Cx{tap.re, -tap.im});
  st.ytf = cdiv(Cx{-ys.re, -ys.im}, tap);
  return st;
}

int YBus::find(int row, int col) const {
  auto b = ci.begin() + rp[row], e = ci.begin() + rp[row + 1];
  auto it = std::lower_bound(b, e, col);
  return (it != e && *it == col) ? (int)(it - ci.begin()) : -1;
}

Network copy_network(const gs_network& s) {
  Network n;
  n.n = s.n_bus;
  n.nb = s.n_branch;
  n.ng = s.n_gen;
  n.base = s.base_mva;
  auto cp = [](auto& dst, const auto* src, int cnt) { dst.assign(src, src + cnt); };
  cp(n.bus_type, s.bus_type, n.n);
  cp(n.pd, s.pd, n.n); cp(n.qd, s.qd, n.n); cp(n.gs, s.gs, n.n); cp(n.bs, s.bs, n.n);
  cp(n.vm, s.vm, n.n); cp(n.va_deg, s.va_deg, n.n);
  cp(n.br_from, s.br_from, n.nb); cp(n.br_to, s.br_to, n.nb);
  cp(n.br_r, s.br_r, n.nb); cp(n.br_x, s.br_x, n.nb); cp(n.br_b, s.br_b, n.nb);
  cp(n.br_tap, s.br_tap, n.nb); cp(n.br_shift, s.br_shift_deg, n.nb); cp(n.br_in, s.br_in_service, n.nb);
  cp(n.gen_bus, s.gen_bus, n.ng); cp(n.gen_pg, s.gen_pg, n.ng); cp(n.gen_qg, s.gen_qg, n.ng);
  cp(n.gen_vg, s.gen_vg, n.ng);
  const double nan = std::nan("");
  if (s.br_rate_a) cp(n.br_rate, s.br_rate_a, n.nb); else n.br_rate.assign(n.nb, 0.0);
  if (s.bus_vmin) cp(n.bus_lo, s.bus_vmin, n.n); else n.bus_lo.assign(n.n, nan);
  if (s.bus_vmax) cp(n.bus_hi, s.bus_vmax, n.n); else n.bus_hi.assign(n.n, nan);
  if (s.gen_qmin) cp(n.gen_qmin, s.gen_qmin, n.ng); else n.gen_qmin.assign(n.ng, -INFINITY);
  if (s.gen_qmax) cp(n.gen_qmax, s.gen_qmax, n.ng); else n.gen_qmax.assign(n.ng, INFINITY);
  return n;
}

YBus build_ybus(const Network& net, int skip) {
  struct Trip { int row, col, branch; double g, b; };
  vector<Trip> trips;
  trips.reserve(4 * (size_t)net.nb + net.n);
  auto add_trip = [&](int r, int c, int br, double g, double b) {
    if (g == 0.0 && b == 0.0) return;  // ComplexCsr.Builder.Add skips an all-zero stamp
    trips.push_back({r, c, br, g, b});
  };
  for (int k = 0; k < net.nb; k++) {
    if (k == skip || !net.br_in[k]) continue;
    int f = net.br_from[k], t = net.br_to[k];
    Stamp s = branch_stamp(net.br_r[k], net.br_x[k], net.br_b[k], net.br_tap[k], net.br_shift[k]);
    add_trip(f, f, k, s.yff.re, s.yff.im);
    add_trip(t, t, k, s.ytt.re, s.ytt.im);
    add_trip(f, t, k, s.yft.re, s.yft.im);
    add_trip(t, f, k, s.ytf.re, s.ytf.im);
  }
  for (int i = 0; i < net.n; i++) {
    if (net.gs[i] == 0 && net.bs[i] == 0) continue;
    add_trip(i, i, -1, net.gs[i] / net.base, net.bs[i] / net.base);
  }
  // Bucket by row keeping insertion order, then a stable sort by column inside each row.
  int n = net.n;
  vector<int> start(n + 1, 0);
  for (auto& t : trips) start[t.row + 1]++;
  for (int i = 0; i < n; i++) start[i + 1] += start[i];
  vector<int> order(trips.size()), fill(start.begin(), start.end() - 1);
  for (size_t t = 0; t < trips.size(); t++) order[fill[trips[t].row]++] = (int)t;
  YBus y;
  y.n = n;
  y.rp.assign(n + 1, 0);
  y.diag.assign(n, -1);
  y.contrib_ptr.push_back(0);
  auto open_cell = [&](int col) {
    y.ci.push_back(col);
    y.g.push_back(0.0);
    y.b.push_back(0.0);
    y.contrib_ptr.push_back(y.contrib_ptr.back());
  };
  for (int i = 0; i < n; i++) {
    std::stable_sort(order.begin() + start[i], order.begin() + start[i + 1],
                     [&](int a, int b) { return trips[a].col < trips[b].col; });
    y.rp[i] = (int)y.ci.size();
    int prev = std::numeric_limits<int>::min();
    for (int p = start[i]; p < start[i + 1]; p++) {
      const Trip& t = trips[order[p]];
      if (t.col != prev) {
        if (y.diag[i] < 0 && prev < i && t.col > i) { y.diag[i] = (int)y.ci.size(); open_cell(i); }
        open_cell(t.col);
        if (t.col == i) y.diag[i] = (int)y.ci.size() - 1;
        prev = t.col;
      }
      // duplicates accumulate left to right from +0.0 (0.0 + first stamp, then += each further stamp)
      y.g.back() += t.g;
      y.b.back() += t.b;
      y.contrib_branch.push_back(t.branch);
      y.contrib_g.push_back(t.g);
      y.contrib_b.push_back(t.b);
      y.contrib_ptr.back()++;
    }
    if (y.diag[i] < 0) { y.diag[i] = (int)y.ci.size(); open_cell(i); }
  }
  y.rp[n] = (int)y.ci.size();
  return y;
}

BusSpec build_bus_spec(const Network& net, bool flat_start, double load_scale) {
  int n = net.n;
  BusSpec s;
  s.type.assign(n, GS_BUS_PQ);
  s.vm.assign(n, 0.0); s.va.assign(n, 0.0); s.pspec.assign(n, 0.0); s.qspec.assign(n, 0.0);
  vector<double> pg(n, 0.0), qg(n, 0.0), vg(n, 0.0);
  vector<char> has(n, 0);
  s.qmin.assign(n, 0.0);
  s.qmax.assign(n, 0.0);
  for (int g = 0; g < net.ng; g++) {
    int i = net.gen_bus[g];
    pg[i] += net.gen_pg[g] * load_scale;
    qg[i] += net.gen_qg[g];
    s.qmin[i] += g < (int)net.gen_qmin.size() ? net.gen_qmin[g] : -INFINITY;
    s.qmax[i] += g < (int)net.gen_qmax.size() ? net.gen_qmax[g] : INFINITY;
    vg[i] = net.gen_vg[g];
    has[i] = 1;
  }
  s.vg = vg;
  s.qd.resize(n);
  for (int i = 0; i < n; i++) s.qd[i] = net.qd[i] * load_scale;
  for (int i = 0; i < n; i++) {
    int bt = net.bus_type[i];
    s.type[i] = (bt == GS_BUS_PV && !has[i]) ? GS_BUS_PQ : bt;
    bool vc = s.type[i] == GS_BUS_SLACK || s.type[i] == GS_BUS_PV;
    s.vm[i] = vc && has[i] ? vg[i] : s.type[i] == GS_BUS_SLACK ? net.vm[i] : flat_start ? 1.0 : net.vm[i];
    s.va[i] = bt == GS_BUS_SLACK ? net.va_deg[i] * kPi / 180.0 : flat_start ? 0.0 : net.va_deg[i] * kPi / 180.0;
    s.pspec[i] = (pg[i] - net.pd[i] * load_scale) / net.base;
    s.qspec[i] = (qg[i] - net.qd[i] * load_scale) / net.base;
  }
  return s;
}

JacobianPattern build_jacobian_pattern(const YBus& y, const vector<int>& type) {
  JacobianPattern jp;
  int n = y.n;
  jp.ppos.assign(n, -1);
  jp.qpos.assign(n, -1);
  int m = 0;
  for (int i = 0; i < n; i++) if (type[i] != GS_BUS_SLACK) jp.ppos[i] = m++;
  for (int i = 0; i < n; i++) if (type[i] == GS_BUS_PQ) jp.qpos[i] = m++;
  jp.m = m;
  // Cells per Jacobian row: every (Y cell, block) maps to a distinct cell, so no duplicates arise; sort by column.
  vector<vector<int>> rows(m);
  for (int i = 0; i < n; i++) {
    int pi = jp.ppos[i], qi = jp.qpos[i];
    if (pi < 0 && qi < 0) continue;
    for (int k = y.rp[i]; k < y.rp[i + 1]; k++) {
      int j = y.ci[k], pj = jp.ppos[j], qj = jp.qpos[j];
      if (pi >= 0 && pj >= 0) rows[pi].push_back(pj);
      if (pi >= 0 && qj >= 0) rows[pi].push_back(qj);
      if (qi >= 0 && pj >= 0) rows[qi].push_back(pj);
      if (qi >= 0 && qj >= 0) rows[qi].push_back(qj);
    }
  }
  jp.rp.assign(m + 1, 0);
  for (int r = 0; r < m; r++) {
    std::sort(rows[r].begin(), rows[r].end());
    rows[r].erase(std::unique(rows[r].begin(), rows[r].end()), rows[r].end());
    jp.rp[r + 1] = jp.rp[r] + (int)rows[r].size();
  }
  jp.ci.reserve(jp.rp[m]);
  for (int r = 0; r < m; r++) jp.ci.insert(jp.ci.end(), rows[r].begin(), rows[r].end());
  auto locate = [&](int r, int c) {
    auto b = jp.ci.begin() + jp.rp[r], e = jp.ci.begin() + jp.rp[r + 1];
    return (int)(std::lower_bound(b, e, c) - jp.ci.begin());
  };
  jp.slots.assign(4 * (size_t)y.rp[n], -1);
  for (int i = 0; i < n; i++) {
    int pi = jp.ppos[i], qi = jp.qpos[i];
    for (int k = y.rp[i]; k < y.rp[i + 1]; k++) {
      int j = y.ci[k], pj = jp.ppos[j], qj = jp.qpos[j];
      size_t b = 4 * (size_t)k;
      if (pi >= 0 && pj >= 0) jp.slots[b + 0] = locate(pi, pj);
      if (pi >= 0 && qj >= 0) jp.slots[b + 1] = locate(pi, qj);
      if (qi >= 0 && pj >= 0) jp.slots[b + 2] = locate(qi, pj);
      if (qi >= 0 && qj >= 0) jp.slots[b + 3] = locate(qi, qj);
    }
  }
  return jp;
}

void host_injections(const YBus& y, const double* vm, const double* va, double* p, double* q) {
  for (int i = 0; i < y.n; i++) {
    double pi = 0, qi = 0, vai = va[i];
    for (int k = y.rp[i]; k < y.rp[i + 1]; k++) {
      int j = y.ci[k];
      double sn, cs;
      gs_fdlibm::sincos(vai - va[j], &sn, &cs);
      double vmj = vm[j];
      pi += vmj * (y.g[k] * cs + y.b[k] * sn);
      qi += vmj * (y.g[k] * sn - y.b[k] * cs);
    }
    p[i] = vm[i] * pi;
    q[i] = vm[i] * qi;
  }
}

void host_jacobian(const YBus& y, const JacobianPattern& jp, const double* vm, const double* va, const double* p,
                   const double* q, double* val) {
  std::fill(val, val + jp.rp[jp.m], 0.0);
  for (int i = 0; i < y.n; i++) {
    for (int k = y.rp[i]; k < y.rp[i + 1]; k++) {
      int j = y.ci[k];
      double g = y.g[k], b = y.b[k], h, nn, mm, l;
      if (i == j) {
        h = -q[i] - b * vm[i] * vm[i];
        nn = p[i] / vm[i] + g * vm[i];
        mm = p[i] - g * vm[i] * vm[i];
        l = q[i] / vm[i] - b * vm[i];
      } else {
        double sn, cs;
        gs_fdlibm::sincos(va[i] - va[j], &sn, &cs);
        h = vm[i] * vm[j] * (g * sn - b * cs);
        nn = vm[i] * (g * cs + b * sn);
        mm = -vm[i] * vm[j] * (g * cs + b * sn);
        l = vm[i] * (g * sn - b * cs);
      }
      const int* s = &jp.slots[4 * (size_t)k];
      if (s[0] >= 0) val[s[0]] = h;
      if (s[1] >= 0) val[s[1]] = nn;
      if (s[2] >= 0) val[s[2]] = mm;
      if (s[3] >= 0) val[s[3]] = l;
    }
  }
}

// Exact minimum degree on the explicit elimination graph with a bucket queue: LIFO among equal degrees, first
// insertion in descending index (the lowest index wins the first ties), neighbours reinserted in adjacency order.
vector<int> minimum_degree(int n, const int* rowptr, const int* colidx) {
  vector<int> order(n);
  if (n == 0) return order;
  vector<vector<int>> adj(n);
  {
    vector<int> cnt(n, 0);
    for (int i = 0; i < n; i++)
      for (int k = rowptr[i]; k < rowptr[i + 1]; k++)
        if (colidx[k] != i) { cnt[i]++; cnt[colidx[k]]++; }
    for (int i = 0; i < n; i++) adj[i].reserve(cnt[i]);
    for (int i = 0; i < n; i++)
      for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
        int j = colidx[k];
        if (j == i) continue;
        adj[i].push_back(j);
        adj[j].push_back(i);
      }
  }
  vector<int> mark(n, 0);
  int tick = 0;
  for (int i = 0; i < n; i++) {  // de-duplicate keeping first occurrences
    tick++;
    auto& a = adj[i];
    size_t w = 0;
    for (size_t p = 0; p < a.size(); p++)
      if (mark[a[p]] != tick) { mark[a[p]] = tick; a[w++] = a[p]; }
    a.resize(w);
  }
  vector<int> head(n + 1, -1), nxt(n, -1), prv(n, -1), deg(n, 0);
  auto push = [&](int v, int d) {
    deg[v] = d; prv[v] = -1; nxt[v] = head[d];
    if (head[d] >= 0) prv[head[d]] = v;
    head[d] = v;
  };
  auto pop = [&](int v) {
    if (prv[v] >= 0) nxt[prv[v]] = nxt[v]; else head[deg[v]] = nxt[v];
    if (nxt[v] >= 0) prv[nxt[v]] = prv[v];
  };
  for (int i = n - 1; i >= 0; i--) push(i, (int)adj[i].size());
  vector<char> gone(n, 0);
  vector<int> nb;
  nb.reserve(n);
  int lo = 0;
  for (int s = 0; s < n; s++) {
    while (head[lo] < 0) lo++;
    int v = head[lo];
    pop(v);
    gone[v] = 1;
    order[s] = v;
    nb.clear();
    for (int u : adj[v]) if (!gone[u]) nb.push_back(u);
    tick++;
    for (int u : nb) mark[u] = tick;
    for (int u : nb) {
      pop(u);
      auto& au = adj[u];
      size_t w = 0;
      for (size_t p = 0; p < au.size(); p++) {
        int x = au[p];
        if (x != v && !gone[x] && mark[x] != tick) au[w++] = x;
      }
      au.resize(w);
      for (int x : nb) if (x != u) au.push_back(x);
      push(u, (int)au.size());
    }
    adj[v].clear();
    adj[v].shrink_to_fit();
    if (lo > 0) lo--;
  }
  return order;
}

bool gp_plan(int m, const vector<int>& rowptr, const vector<int>& colidx, const double* val, const vector<int>& q,
             double diag_pref, bool keep_zeros, Plan& plan, std::string& why) {
  plan = Plan();
  plan.m = m;
  plan.diagonal_preference = diag_pref;
  // CSC of A: every structural cell (ToCscKeepingZeros) or the non-zero ones (ToCsc); a_src maps back to CSR.
  plan.ap.assign(m + 1, 0);
  int stored = rowptr[m];
  for (int k = 0; k < stored; k++)
    if (keep_zeros || val[k] != 0.0) plan.ap[colidx[k] + 1]++;
  for (int c = 0; c < m; c++) plan.ap[c + 1] += plan.ap[c];
  int nnz = plan.ap[m];
  plan.ai.assign(nnz, 0);
  plan.a_src.assign(nnz, 0);
  vector<double> ax(nnz);
  {
    vector<int> fill(plan.ap.begin(), plan.ap.end() - 1);
    for (int i = 0; i < m; i++)
      for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
        if (!keep_zeros && val[k] == 0.0) continue;
        int d = fill[colidx[k]]++;
        plan.ai[d] = i;
        plan.a_src[d] = k;
        ax[d] = val[k];
      }
  }
  // row floors: 1e-12 x the largest |entry| of each original row; NaN / Inf refused
  vector<double> floor_row(m, 0.0);
  for (int e = 0; e < nnz; e++) {
    double v = ax[e];
    if (!std::isfinite(v)) { why = "the Jacobian holds a non-finite entry"; return false; }
    double a = std::fabs(v);
    if (a > floor_row[plan.ai[e]]) floor_row[plan.ai[e]] = a;
  }
  for (double& f : floor_row) f *= 1e-12;
  plan.q = q;
  plan.pinv.assign(m, -1);
  plan.pivot_row.assign(m, -1);
  plan.lp.assign(m + 1, 0);
  plan.up.assign(m + 1, 0);
  plan.reach_ptr.assign(m + 1, 0);
  vector<double> x(m, 0.0);
  vector<int> xi(m), stack_pos(m);
  vector<char> seen(m, 0);
  vector<int>& li = plan.li_orig;
  vector<double> lx;
  for (int k = 0; k < m; k++) {
    plan.lp[k] = (int)li.size();
    plan.up[k] = (int)plan.ui.size();
    int col = q[k];
    // reach of A(:,col) in the graph of L, by iterative depth-first search; xi[top..m) in topological order
    int top = m;
    for (int p = plan.ap[col]; p < plan.ap[col + 1]; p++) {
      int start = plan.ai[p];
      if (seen[start]) continue;
      int head = 0;
      xi[0] = start;
      while (head >= 0) {
        int j = xi[head];
        int jcol = plan.pinv[j];
        if (!seen[j]) { seen[j] = 1; stack_pos[head] = jcol < 0 ? 0 : plan.lp[jcol] + 1; }
        bool done = true;
        int end = jcol < 0 ? 0 : plan.lp[jcol + 1];
        for (int r = stack_pos[head]; r < end; r++) {
          int i = li[r];
          if (seen[i]) continue;
          stack_pos[head] = r + 1;
          xi[++head] = i;
          done = false;
          break;
        }
        if (done) { head--; xi[--top] = j; }
      }
    }
    for (int p = top; p < m; p++) seen[xi[p]] = 0;
    // numeric column: scatter A(:,col), update in topological order
    for (int p = top; p < m; p++) x[xi[p]] = 0.0;
    for (int p = plan.ap[col]; p < plan.ap[col + 1]; p++) x[plan.ai[p]] = ax[p];
    for (int p = top; p < m; p++) {
      int j = xi[p], jcol = plan.pinv[j];
      if (jcol < 0) continue;
      double xj = x[j];
      for (int r = plan.lp[jcol] + 1; r < plan.lp[jcol + 1]; r++) x[li[r]] -= lx[r] * xj;
    }
    plan.reach.insert(plan.reach.end(), xi.begin() + top, xi.begin() + m);
    plan.reach_ptr[k + 1] = (int)plan.reach.size();
    int ipiv = -1;
    double best = -1.0;
    for (int p = top; p < m; p++) {
      int i = xi[p];
      if (plan.pinv[i] < 0) {
        double t = std::fabs(x[i]);
        if (t > best) { best = t; ipiv = i; }
      } else {
        plan.ui.push_back(plan.pinv[i]);
        plan.ux.push_back(x[i]);
      }
    }
    if (ipiv < 0 || best <= floor_row[ipiv]) {
      why = "singular Jacobian: no pivot above its row floor at step " + std::to_string(k);
      return false;
    }
    if (diag_pref > 0.0 && !(ipiv == col || plan.pinv[col] >= 0)) {
      for (int p = top; p < m; p++) {
        if (xi[p] != col) continue;
        double td = std::fabs(x[col]);
        if (td >= diag_pref * best && td > floor_row[col]) ipiv = col;
        break;
      }
    }
    double pivot = x[ipiv];
    plan.pivot_row[k] = ipiv;
    plan.ui.push_back(k);
    plan.ux.push_back(pivot);
    plan.pinv[ipiv] = k;
    li.push_back(ipiv);
    lx.push_back(1.0);
    for (int p = top; p < m; p++) {
      int i = xi[p];
      if (plan.pinv[i] < 0) { li.push_back(i); lx.push_back(x[i] / pivot); }
      x[i] = 0.0;
    }
  }
  plan.lp[m] = (int)li.size();
  plan.up[m] = (int)plan.ui.size();
  plan.li.resize(li.size());
  for (size_t p = 0; p < li.size(); p++) plan.li[p] = plan.pinv[li[p]];
  plan.lx = std::move(lx);
  return true;
}

// SparseLuFactorization.Solve (SparseLuFactorization.cs:338-371) on a plan's own factors: x[pinv[k]] = b[k], the
// forward column sweep over unit-diagonal L, the back column sweep over U (each column's pivot last), out[q[k]] = x[k].
void plan_solve(const Plan& f, const double* b, double* out) {
  const int n = f.m;
  vector<double> x(n);
  for (int k = 0; k < n; k++) x[f.pinv[k]] = b[k];
  for (int j = 0; j < n; j++) {
    const double xj = x[j];
    for (int p = f.lp[j] + 1; p < f.lp[j + 1]; p++) x[f.li[p]] -= f.lx[p] * xj;
  }
  for (int j = n - 1; j >= 0; j--) {
    const double xj = x[j] / f.ux[f.up[j + 1] - 1];
    x[j] = xj;
    for (int p = f.up[j]; p < f.up[j + 1] - 1; p++) x[f.ui[p]] -= f.ux[p] * xj;
  }
  for (int k = 0; k < n; k++) out[f.q[k]] = x[k];
}

// NewtonShell.TrySeedDcAngles (NewtonShell.cs:282-299) over DcPowerFlow.Solve(DcDemand.LegacyView(model))
// (DcPowerFlow.cs:78-86, DcDemand.cs:38-50, DistributionFactors.cs:143-360,454-520). The reactance-only view: every
// tap 1, every shift 0, every Gs 0, so an in-service branch stamps b = 1 / x into the reduced nodal susceptance B'
// (branch order; the reference bus's row and column left out), the right-hand side is each bus's net injection
// (0 - Pd, then + Pg per generator in generator order) / base, B' is factored as SparseLu.Factor does (minimum degree,
// explicit zeros dropped, partial pivoting) and solved; each angle is reduced to [-180, 180] degrees (the IEEE
// remainder Math.IEEERemainder computes) and offset by the slack's start angle. False, va untouched, unless exactly one
// bus is a slack, the in-service branches (none of zero reactance) connect every bus, and B' factors.
bool dc_seed_angles(const Network& net, const vector<int>& type, vector<double>& va) {
  const int n = net.n;
  int slack = -1, slacks = 0;
  for (int i = 0; i < n; i++)
    if (type[i] == GS_BUS_SLACK) { slack = i; slacks++; }
  if (slacks != 1) return false;
  vector<int> parent(n);
  for (int i = 0; i < n; i++) parent[i] = i;
  auto find = [&](int v) { while (parent[v] != v) { parent[v] = parent[parent[v]]; v = parent[v]; } return v; };
  for (int l = 0; l < net.nb; l++) {
    if (!net.br_in[l]) continue;
    if (net.br_x[l] == 0.0) return false;   // the reactance-only model cannot take a zero-reactance branch
    const int a = find(net.br_from[l]), c = find(net.br_to[l]);
    if (a != c) parent[a] = c;
  }
  for (int i = 0; i < n; i++)
    if (find(i) != find(slack)) return false;   // more than one DC island
  const int nr = n - 1;
  if (nr == 0) return false;
  vector<int> red(n);
  for (int i = 0, r = 0; i < n; i++) red[i] = i == slack ? -1 : r++;
  // B' by stamps in branch order, duplicates summed in insertion order from +0.0, an explicit diagonal in every row
  // (CsrMatrix.Builder), then the columns sorted per row with a stable sort
  vector<vector<std::pair<int, double>>> rows(nr);
  auto add = [&](int r, int c, double v) { rows[r].push_back({c, v}); };
  for (int l = 0; l < net.nb; l++) {
    if (!net.br_in[l]) continue;
    const int f = red[net.br_from[l]], t = red[net.br_to[l]];
    const double bl = 1.0 / net.br_x[l];
    if (f >= 0) add(f, f, bl);
    if (t >= 0) add(t, t, bl);
    if (f >= 0 && t >= 0) { add(f, t, -bl); add(t, f, -bl); }
  }
  vector<int> rp(nr + 1, 0), ci;
  vector<double> val;
  for (int r = 0; r < nr; r++) {
    auto& row = rows[r];
    std::stable_sort(row.begin(), row.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    bool diag = false;
    for (size_t e = 0; e < row.size(); e++) {
      if (!ci.empty() && (int)ci.size() > rp[r] && ci.back() == row[e].first) { val.back() += row[e].second; continue; }
      if (!diag && row[e].first > r) { ci.push_back(r); val.push_back(0.0); diag = true; }
      ci.push_back(row[e].first);
      val.push_back(row[e].second);
      diag |= row[e].first == r;
    }
    if (!diag) { ci.push_back(r); val.push_back(0.0); }
    rp[r + 1] = (int)ci.size();
  }
  Plan f;
  std::string why;
  const vector<int> order = minimum_degree(nr, rp.data(), ci.data());
  if (!gp_plan(nr, rp, ci, val.data(), order, 0.0, false, f, why)) return false;
  vector<double> inj(n), rhs(nr, 0.0), x(nr);
  for (int i = 0; i < n; i++) inj[i] = 0.0 - net.pd[i];
  for (int g = 0; g < net.ng; g++) inj[net.gen_bus[g]] += net.gen_pg[g];
  for (int i = 0; i < n; i++)
    if (red[i] >= 0) rhs[red[i]] += inj[i] / net.base;
  plan_solve(f, rhs.data(), x.data());
  const double va0 = va[slack];
  for (int i = 0; i < n; i++) {
    if (i == slack) continue;
    const double deg = x[red[i]] * 180.0 / kPi;
    va[i] = va0 + std::remainder(deg, 360.0) * kPi / 180.0;
  }
  return true;
}

// NewtonShell.MaxBranchAngleDeg (NewtonShell.cs:311-320): the largest angle across the series element of any in-service
// branch, net of its phase shift, wrapped to [-180, 180] degrees.
double max_branch_angle_deg(const Network& net, const double* va) {
  double worst = 0;
  for (int l = 0; l < net.nb; l++) {
    if (!net.br_in[l]) continue;
    const double df = va[net.br_from[l]] * 180.0 / kPi, dt = va[net.br_to[l]] * 180.0 / kPi;
    const double d = std::fabs(std::remainder(df - net.br_shift[l] - dt, 360.0));
    if (d > worst) worst = d;
  }
  return worst;
}

vector<int> symmetric_cell_transpose(const YBus& y) {
  const int nnz = y.rp[y.n];
  vector<int> tr(nnz, -1);
  auto same = [](double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; };
  for (int i = 0; i < y.n; i++)
    for (int k = y.rp[i]; k < y.rp[i + 1]; k++) {
      const int t = y.find(y.ci[k], i);
      if (t < 0 || !same(y.g[k], y.g[t]) || !same(y.b[k], y.b[t])) return {};
      const int c0 = y.contrib_ptr[k], c1 = y.contrib_ptr[k + 1], d0 = y.contrib_ptr[t];
      if (c1 - c0 != y.contrib_ptr[t + 1] - d0) return {};
      for (int c = 0; c < c1 - c0; c++)
        if (y.contrib_branch[c0 + c] != y.contrib_branch[d0 + c] || !same(y.contrib_g[c0 + c], y.contrib_g[d0 + c]) ||
            !same(y.contrib_b[c0 + c], y.contrib_b[d0 + c]))
          return {};
      tr[k] = t;
    }
  return tr;
}

int jacobian_value_positions(const YBus& y, const JacobianPattern& jp, const vector<int>& tr, vector<int>& pos) {
  const int nnz = y.rp[y.n];
  pos.assign(jp.rp[jp.m], -1);
  for (int i = 0; i < y.n; i++)
    for (int k = y.rp[i]; k < y.rp[i + 1]; k++) {
      // the lane of cell tr[k] forms cell k's values (tr is its own inverse; a diagonal cell is its own transpose)
      const int base = 4 * (tr.empty() ? k : tr[k]);
      for (int b = 0; b < 4; b++) {
        const int s = jp.slots[4 * (size_t)k + b];
        if (s >= 0) pos[s] = base + b;
      }
    }
  for (int p : pos)
    if (p < 0) throw std::runtime_error("plan: a Jacobian entry without a Y cell");
  return 4 * nnz;
}

void compile_plan(const Plan& plan, DevicePlanHost& d, int scratch_cap, const vector<int>* a_pos, int a_slots) {
  int m = plan.m;
  d = DevicePlanHost();
  d.scratch_cap = scratch_cap;
  d.m = m;
  d.lp = plan.lp;
  d.up = plan.up;
  d.li = plan.li;
  d.ui = plan.ui;
  d.pinv = plan.pinv;
  d.q = plan.q;
  d.nnz_l = plan.lp[m];
  d.nnz_u = plan.up[m];
  // Column k's workspace: slots [0, nU_k - 1) the off-diagonal U rows in storage (= topological) order, slot nU_k - 1
  // the pivot, slots [nU_k, nU_k + nL_k - 1) the off-diagonal L rows in storage order.
  vector<int> slot_of(m, -1);
  d.col_a_ptr.assign(m + 1, 0);
  d.pivot_slot.assign(m, 0);
  d.upd_ptr.assign(d.nnz_u + 1, 0);
  for (int k = 0; k < m; k++) {
    int nu = plan.up[k + 1] - plan.up[k], nl = plan.lp[k + 1] - plan.lp[k];
    d.max_col = std::max(d.max_col, nu + nl - 1);
    for (int t = 0; t < nu - 1; t++) slot_of[plan.ui[plan.up[k] + t]] = t;
    slot_of[k] = nu - 1;
    d.pivot_slot[k] = nu - 1;
    for (int t = 1; t < nl; t++) slot_of[plan.li[plan.lp[k] + t]] = nu - 1 + t;
    int col = plan.q[k];
    for (int e = plan.ap[col]; e < plan.ap[col + 1]; e++) {
      int s = slot_of[plan.pinv[plan.ai[e]]];
      if (s < 0) throw std::runtime_error("plan: an entry of A lies outside its column's pattern");
      d.a_src.push_back(plan.a_src[e]);
      d.a_slot.push_back(s);
    }
    d.col_a_ptr[k + 1] = (int)d.a_src.size();
    for (int t = 0; t < nu; t++) {
      int ue = plan.up[k] + t;
      d.upd_ptr[ue] = (int)d.upd_slot.size();
      if (t < nu - 1) {
        int j = plan.ui[ue];
        for (int r = plan.lp[j] + 1; r < plan.lp[j + 1]; r++) {
          int s = slot_of[plan.li[r]];
          if (s < 0) throw std::runtime_error("plan: an update lies outside its column's pattern");
          d.upd_slot.push_back(s);
        }
      }
    }
    for (int t = 0; t < nu - 1; t++) slot_of[plan.ui[plan.up[k] + t]] = -1;
    slot_of[k] = -1;
    for (int t = 1; t < nl; t++) slot_of[plan.li[plan.lp[k] + t]] = -1;
  }
  d.upd_ptr[d.nnz_u] = (int)d.upd_slot.size();
  d.n_upd = (int)d.upd_slot.size();
  d.prow = plan.pivot_row;
  // Dependency levels: column k reads L(:, j) for every off-diagonal U(j, k), so level(k) = 1 + max level(j).
  vector<int> level(m, 0);
  int n_levels = 0;
  for (int k = 0; k < m; k++) {
    int lv = 0;
    for (int p = plan.up[k]; p < plan.up[k + 1] - 1; p++) lv = std::max(lv, level[plan.ui[p]] + 1);
    level[k] = lv;
    n_levels = std::max(n_levels, lv + 1);
  }
  vector<int> count(n_levels + 1, 0);
  for (int k = 0; k < m; k++) count[level[k] + 1]++;
  for (int l = 0; l < n_levels; l++) count[l + 1] += count[l];
  d.level_end.assign(count.begin() + 1, count.end());
  d.level_end.push_back(m + 1);  // sentinel past the last level
  d.col_order.assign(m, 0);
  for (int k = 0; k < m; k++) d.col_order[count[level[k]]++] = k;
  d.n_levels = n_levels;
  // Level chunks: within each level, columns by decreasing chain length, kTracks at a time.
  const int G = DevicePlanHost::kTracks;
  // An odd stride puts the same slot of the 8 tracks in 8 different LDS banks (the tracks' updates tend to touch the
  // same slot positions at once).
  d.wstride = d.max_col | 1;
  if (d.wstride * G + DevicePlanHost::kSpareSlots > 0xFFFF)
    throw std::runtime_error("plan: column workspace too large for 16-bit slot indices");
  d.ch_track_ptr.assign(1, 0);
  d.ch_row_ptr.assign(1, 0);
  d.ch_a_ptr.assign(1, 0);
  d.ch_tl.clear();
  for (int lv = 0; lv < n_levels; lv++) {
    int b0 = lv == 0 ? 0 : d.level_end[lv - 1], b1 = d.level_end[lv];
    vector<int> cols(d.col_order.begin() + b0, d.col_order.begin() + b1);
    std::stable_sort(cols.begin(), cols.end(), [&](int a, int b) {
      return (plan.up[a + 1] - plan.up[a]) > (plan.up[b + 1] - plan.up[b]);
    });
    // a level of fewer than kTracks columns gives each column more lanes: 64 for one, 32 for two, 16 for four
    int gl = G;
    while (gl > 1 && gl / 2 >= (int)cols.size()) gl /= 2;
    const int TL = 64 / gl;
    for (size_t c0 = 0; c0 < cols.size(); c0 += gl) {
      size_t c1 = std::min(cols.size(), c0 + gl);
      d.ch_tl.push_back(TL);
      // Rows of each track: its column's updates list-scheduled, up to TL per row. An update x[dst] -= L * x[src]
      // (src: the slot of U entry t, dst: an L row of that entry's column) goes in the first row with a free lane
      // after every update of its source slot (x[src] is then final) and after the previous update of its
      // destination slot (each slot receives its updates in the reference's order, one row at a time). The
      // reference's order reaches a source slot's updates before the slot is read, so one pass suffices.
      struct Upd { int rec, lsrc; };
      vector<vector<vector<Upd>>> trows(c1 - c0);  // per track: rows, each up to TL updates
      size_t nrows = 0;
      for (size_t g = 0; g < c1 - c0; g++) {
        int k = cols[c0 + g];
        d.tr_col.push_back(k);
        int nu = plan.up[k + 1] - plan.up[k];
        const int base = (int)g * d.wstride;
        vector<int> final_row(d.wstride, 0), next_row(d.wstride, 0);  // first row a slot may be read / next written
        auto& rows_g = trows[g];
        for (int t = 0; t < nu - 1; t++) {
          int ue = plan.up[k] + t, r0 = d.upd_ptr[ue], r1 = d.upd_ptr[ue + 1];
          int j = plan.ui[ue];
          for (int r = r0; r < r1; r++) {
            const int dst = d.upd_slot[r];
            size_t row = (size_t)std::max(final_row[t], next_row[dst]);
            while (row < rows_g.size() && (int)rows_g[row].size() >= TL) row++;
            if (row == rows_g.size()) rows_g.emplace_back();
            rows_g[row].push_back({(base + dst) | ((base + t) << 16), plan.lp[j] + 1 + (r - d.upd_ptr[ue])});
            next_row[dst] = (int)row + 1;
            final_row[dst] = std::max(final_row[dst], (int)row + 1);
          }
        }
        nrows = std::max(nrows, rows_g.size());
        int col = plan.q[k];
        for (int e = d.col_a_ptr[k]; e < d.col_a_ptr[k + 1]; e++) {
          d.ch_a_ws.push_back((int)g * d.wstride + d.a_slot[e]);
          d.ch_a_src.push_back(d.a_src[e]);
        }
        (void)col;
      }
      for (size_t g = 0; g < c1 - c0; g++) {
        int k = d.tr_col[d.tr_col.size() - (c1 - c0) + g];
        d.tr_u0.push_back(plan.up[k]);
        d.tr_nu.push_back(plan.up[k + 1] - plan.up[k]);
        d.tr_l0.push_back(plan.lp[k]);
        d.tr_nl.push_back(plan.lp[k + 1] - plan.lp[k]);
      }
      const bool level_last = c1 == cols.size();
      if (nrows == 0) {  // a chunk without updates still gets one (empty) row to carry its events
        for (int lane = 0; lane < 64; lane++) { d.row_rec.push_back(0xFFFF); d.row_lsrc.push_back(0); }
        d.row_flags.push_back(DevicePlanHost::kRowBegin | DevicePlanHost::kRowEnd |
                              (level_last ? DevicePlanHost::kRowLevelEnd : 0));
        d.row_level.push_back(lv);
      }
      for (size_t rho = 0; rho < nrows; rho++) {
        d.row_flags.push_back((rho == 0 ? DevicePlanHost::kRowBegin : 0) |
                              (rho + 1 == nrows ? DevicePlanHost::kRowEnd : 0) |
                              (rho + 1 == nrows && level_last ? DevicePlanHost::kRowLevelEnd : 0));
        d.row_level.push_back(lv);
        for (int lane = 0; lane < 64; lane++) {
          size_t g = lane / TL;
          int idx = lane % TL;
          int rec = 0xFFFF, lsrc = 0;
          if (g < c1 - c0 && rho < trows[g].size() && idx < (int)trows[g][rho].size()) {
            rec = trows[g][rho][idx].rec;
            lsrc = trows[g][rho][idx].lsrc;
          }
          d.row_rec.push_back(rec);
          d.row_lsrc.push_back(lsrc);
        }
      }
      d.ch_track_ptr.push_back((int)d.tr_col.size());
      d.ch_row_ptr.push_back(d.ch_row_ptr.back() + (int)std::max<size_t>(nrows, 1));
      d.ch_a_ptr.push_back((int)d.ch_a_ws.size());
      d.ch_fence.push_back(c1 == cols.size() ? 1 : 0);
    }
  }
  d.n_chunks = (int)d.ch_fence.size();
  d.n_rows = d.ch_row_ptr.back();
  // Factor positions: chunk c's w-th WRITE row is the stream's (W_c + w)-th, and its lane g * TL + t % TL stores slot
  // t = w * TL + t % TL of track g, at 64 * (W_c + w) + lane of the factor array.
  d.fpos.assign(d.nnz_l, -1);
  d.bpos.assign(d.nnz_u, -1);
  d.n_wrows = 0;
  for (int c = 0; c < d.n_chunks; c++) {
    const int t0 = d.ch_track_ptr[c], ntr = d.ch_track_ptr[c + 1] - t0, TL = d.ch_tl[c];
    int wrows = 0;
    for (int g = 0; g < ntr; g++) {
      const int k = d.tr_col[t0 + g], u0 = plan.up[k], nu = plan.up[k + 1] - u0, l0 = plan.lp[k];
      const int nslot = nu + plan.lp[k + 1] - l0 - 1;
      wrows = std::max(wrows, (nslot + TL - 1) / TL);
      for (int t = 0; t < nslot; t++) {
        const int pos = (d.n_wrows + t / TL) * 64 + g * TL + t % TL;
        if (t < nu - 1) d.bpos[u0 + t] = pos;
        else if (t > nu - 1) d.fpos[l0 + (t - (nu - 1))] = pos;
      }
    }
    d.n_wrows += wrows;
  }
  d.n_fac = 64 * d.n_wrows;
  d.j2a.assign(plan.a_src.empty() ? 0 : (size_t)(*std::max_element(plan.a_src.begin(), plan.a_src.end()) + 1), -1);
  d.a_pos.resize(d.ch_a_src.size());
  for (size_t e = 0; e < d.ch_a_src.size(); e++) {
    const int s = d.ch_a_src[e];
    d.a_pos[e] = a_pos ? a_pos->at(s) : (int)e;
    if (d.a_pos[e] < 0 || (a_pos && d.a_pos[e] >= a_slots)) throw std::runtime_error("plan: A position out of range");
    d.j2a[s] = d.a_pos[e];
  }
  // Forward solve rows: L(i, j) for j < i, columns ascending; level(i) = 1 + max level(j).
  {
    vector<vector<std::pair<int, int>>> rows(m);
    for (int j = 0; j < m; j++)
      for (int p = plan.lp[j] + 1; p < plan.lp[j + 1]; p++) rows[plan.li[p]].push_back({j, p});
    vector<int> lvl(m, 0);
    int nlv = 0;
    d.lr_ptr.assign(m + 1, 0);
    for (int i = 0; i < m; i++) {
      std::sort(rows[i].begin(), rows[i].end());
      int lv = 0;
      for (auto& e : rows[i]) lv = std::max(lv, lvl[e.first] + 1);
      lvl[i] = lv;
      nlv = std::max(nlv, lv + 1);
      d.lr_ptr[i + 1] = d.lr_ptr[i] + (int)rows[i].size();
      for (auto& e : rows[i]) { d.lr_col.push_back(e.first); d.lr_idx.push_back(e.second); }
    }
    vector<int> cnt(nlv + 1, 0);
    for (int i = 0; i < m; i++) cnt[lvl[i] + 1]++;
    for (int l = 0; l < nlv; l++) cnt[l + 1] += cnt[l];
    d.fw_level_end.assign(cnt.begin() + 1, cnt.end());
    d.fw_order.assign(m, 0);
    for (int i = 0; i < m; i++) d.fw_order[cnt[lvl[i]]++] = i;
    d.fw_levels = nlv;
  }
  // Back solve rows: U(i, j) for j > i, columns descending; level(i) = 1 + max level(j).
  {
    vector<vector<std::pair<int, int>>> rows(m);
    for (int j = 0; j < m; j++)
      for (int p = plan.up[j]; p < plan.up[j + 1] - 1; p++) rows[plan.ui[p]].push_back({j, p});
    vector<int> lvl(m, 0);
    int nlv = 0;
    d.ur_ptr.assign(m + 1, 0);
    for (int i = m - 1; i >= 0; i--) {
      std::sort(rows[i].begin(), rows[i].end(), [](auto& a, auto& b) { return a.first > b.first; });
      int lv = 0;
      for (auto& e : rows[i]) lv = std::max(lv, lvl[e.first] + 1);
      lvl[i] = lv;
      nlv = std::max(nlv, lv + 1);
    }
    for (int i = 0; i < m; i++) {
      d.ur_ptr[i + 1] = d.ur_ptr[i] + (int)rows[i].size();
      for (auto& e : rows[i]) { d.ur_col.push_back(e.first); d.ur_idx.push_back(e.second); }
    }
    vector<int> cnt(nlv + 1, 0);
    for (int i = 0; i < m; i++) cnt[lvl[i] + 1]++;
    for (int l = 0; l < nlv; l++) cnt[l + 1] += cnt[l];
    d.bw_level_end.assign(cnt.begin() + 1, cnt.end());
    d.bw_order.assign(m, 0);
    for (int i = 0; i < m; i++) d.bw_order[cnt[lvl[i]]++] = i;
    d.bw_levels = nlv;
  }
  // Solve step streams: per level, rows by decreasing entry count, 64 rows per pass; per pass (segment) its PRODUCT
  // steps, the entries lane by lane (each lane's entries contiguous in the scratch area, in the reference's order),
  // then one CHAIN step. A PRODUCT lane's value index is its entry's factor position (pos, by CSC index).
  auto build_stream = [&](const vector<int>& order, const vector<int>& level_end, int nlv, const vector<int>& rptr,
                          const vector<int>& rcol, const vector<int>& ridx, bool back, vector<int>& rec,
                          const vector<int>& pos, int& steps, vector<int>& pass_step, vector<int>& level_pass) {
    using D = DevicePlanHost;
    const int cap = d.solve_scratch();
    int b0 = 0;
    steps = 0;
    pass_step.clear();
    level_pass.clear();
    // the step at which each row's value ends in this stream (rows it never writes: long before it starts)
    vector<int> end_step(m, -D::kStepXAhead);
    for (int lv = 0; lv < nlv; lv++) {
      level_pass.push_back((int)pass_step.size());
      int b1 = level_end[lv];
      vector<int> rows;
      for (int t = b0; t < b1; t++) {
        int i = order[t];
        if (back || rptr[i + 1] > rptr[i]) rows.push_back(i);  // forward rows without entries need no work
      }
      std::stable_sort(rows.begin(), rows.end(),
                       [&](int a, int b) { return rptr[a + 1] - rptr[a] > rptr[b + 1] - rptr[b]; });
      for (size_t p0 = 0; p0 < rows.size(); p0 += 64) {
        const int k = (int)std::min<size_t>(64, rows.size() - p0);
        pass_step.push_back(steps);
        vector<int> done(k, 0), len(k);
        for (int l = 0; l < k; l++) len[l] = rptr[rows[p0 + l] + 1] - rptr[rows[p0 + l]];
        vector<char> begun(k, 0), ended(k, 0);
        for (;;) {
          // one segment: up to cap entries, lane by lane
          vector<int> st(k, 0), take(k, 0);
          int fill = 0;
          for (int l = 0; l < k; l++) {
            st[l] = fill;
            take[l] = std::min(len[l] - done[l], cap - fill);
            fill += take[l];
          }
          // PRODUCT steps
          vector<int> ent_col, ent_idx;
          for (int l = 0; l < k; l++)
            for (int e = done[l]; e < done[l] + take[l]; e++) {
              ent_col.push_back(rcol[rptr[rows[p0 + l]] + e]);
              ent_idx.push_back(ridx[rptr[rows[p0 + l]] + e]);
            }
          for (int e0 = 0; e0 < fill; e0 += 64) {
            int fl = D::kStepProduct;
            for (int e = e0; e < std::min(fill, e0 + 64); e++)
              if (end_step[ent_col[e]] > steps - D::kStepXAhead) fl |= D::kProductLate;
            for (int lane = 0; lane < 64; lane++) {
              const int e = e0 + lane;
              int col = -1, idx = 0;
              if (e < fill) {
                col = ent_col[e];
                idx = pos[ent_idx[e]];
                if (idx < 0) throw std::runtime_error("plan: a solve entry without a factor position");
              }
              rec.push_back(col); rec.push_back(idx | (fl << D::kRsMetaShift));
            }
            steps++;
          }
          // the CHAIN step
          for (int lane = 0; lane < 64; lane++) {
            int x0 = 0, row = 0, fl = D::kStepChain | D::kChainIdle;
            if (lane < k && !ended[lane] && (take[lane] > 0 || done[lane] == len[lane])) {
              fl = D::kStepChain;
              row = rows[p0 + lane];
              x0 = st[lane] | (take[lane] << 16);
              if (!begun[lane]) { fl |= D::kChainBegin; begun[lane] = 1; }
              if (done[lane] + take[lane] == len[lane]) {
                fl |= D::kChainEnd;
                ended[lane] = 1;
                end_step[row] = steps;
              }
            }
            rec.push_back(x0); rec.push_back(row | (fl << D::kRsMetaShift));
          }
          steps++;
          bool all = true;
          for (int l = 0; l < k; l++) {
            done[l] += take[l];
            all = all && ended[l];
          }
          if (all) break;
        }
      }
      b0 = b1;
    }
    level_pass.push_back((int)pass_step.size());
    pass_step.push_back(steps);
  };
  build_stream(d.fw_order, d.fw_level_end, d.fw_levels, d.lr_ptr, d.lr_col, d.lr_idx, false, d.fs_rec, d.fpos,
               d.fs_steps, d.fs_pass_step, d.fs_level_pass);
  build_stream(d.bw_order, d.bw_level_end, d.bw_levels, d.ur_ptr, d.ur_col, d.ur_idx, true, d.bs_rec, d.bpos,
               d.bs_steps, d.bs_pass_step, d.bs_level_pass);
  // The refactor's unified row stream (see DevicePlanHost): per chunk its ASSIGN rows, its UPDATE rows (no-op rows
  // first where a row would read a level that ended fewer than kRsValueAhead rows before), a PIVOT row and its WRITE
  // rows. The L values the updates read live in vals at n_a + their factor position.
  {
    using D = DevicePlanHost;
    d.n_a = a_pos ? a_slots : (int)d.ch_a_src.size();
    vector<int> col_of_l(d.nnz_l);
    for (int j = 0; j < m; j++)
      for (int p = plan.lp[j]; p < plan.lp[j + 1]; p++) col_of_l[p] = j;
    vector<int> level_end_row(n_levels, -1);
    auto emit = [&](int kind, int lv, const int* rec2) {
      d.rs_meta.push_back(kind);
      d.rs_level.push_back(lv);
      d.rs_rec.insert(d.rs_rec.end(), rec2, rec2 + 2 * 64);
    };
    if ((int64_t)d.n_a + d.n_fac > D::kRsIndexMask || m > D::kRsIndexMask)
      throw std::runtime_error("plan: too many values for the row stream");
    vector<int> rec(2 * 64);
    auto nop = [&](int lv) {
      std::fill(rec.begin(), rec.end(), 0);
      emit(D::kRsNop, lv, rec.data());
    };
    d.rs_chunk_row.clear();
    d.rs_level_chunk.assign(1, 0);
    int n_wr = 0;   // WRITE rows emitted so far
    for (int c = 0; c < d.n_chunks; c++) {
      const int t0 = d.ch_track_ptr[c], ntr = d.ch_track_ptr[c + 1] - t0;
      const int lv = level[d.tr_col[t0]];
      const int TL = d.ch_tl[c];
      const int width = D::width_code(TL);   // folded into every row of the chunk
      if (lv + 1 < (int)d.rs_level_chunk.size()) throw std::runtime_error("plan: chunks out of level order");
      while ((int)d.rs_level_chunk.size() <= lv) d.rs_level_chunk.push_back(c);
      d.rs_chunk_row.push_back((int)d.rs_meta.size());
      // ASSIGN rows: (slot, A value index)
      const int a0 = d.ch_a_ptr[c], a1 = d.ch_a_ptr[c + 1];
      for (int e0 = a0; e0 < std::max(a1, a0 + 1); e0 += 64) {
        for (int l = 0; l < 64; l++) {
          const int e = e0 + l;
          const bool in = e < a1;
          rec[2 * l] = in ? d.ch_a_ws[e] : d.spare(l);
          rec[2 * l + 1] = in ? d.a_pos[e] : 0;
        }
        emit(D::kRsAssign | (e0 == a0 ? D::kRsClear : 0) | width, lv, rec.data());
      }
      // UPDATE rows: (dst | src << 16, L value index)
      for (int r = d.ch_row_ptr[c]; r < d.ch_row_ptr[c + 1]; r++) {
        int need = -1;
        bool any = false;
        for (int l = 0; l < 64; l++) {
          const int rr = d.row_rec[(size_t)r * 64 + l];
          if ((rr & 0xFFFF) == 0xFFFF) continue;
          any = true;
          need = std::max(need, level[col_of_l[d.row_lsrc[(size_t)r * 64 + l]]]);
        }
        if (!any) continue;
        if (need >= lv || level_end_row[need] < 0) throw std::runtime_error("plan: an update reads an unfinished level");
        while ((int)d.rs_meta.size() < level_end_row[need] + D::kRsValueAhead) nop(lv);
        for (int l = 0; l < 64; l++) {
          const int rr = d.row_rec[(size_t)r * 64 + l];
          const bool in = (rr & 0xFFFF) != 0xFFFF;
          rec[2 * l] = in ? rr : d.spare(l) | (d.spare(l) << 16);
          rec[2 * l + 1] = in ? d.n_a + d.fpos[d.row_lsrc[(size_t)r * 64 + l]] : 0;
        }
        emit(D::kRsUpdate | width, lv, rec.data());
      }
      // PIVOT row: (pivot slot | candidates << 16, pivot step), -1 for an idle track
      int wrows = 0;
      for (int l = 0; l < 64; l++) {
        const int g = l / TL;
        rec[2 * l] = -1;
        rec[2 * l + 1] = 0;
        if (g < ntr) {
          const int k = d.tr_col[t0 + g], nu = plan.up[k + 1] - plan.up[k], nl = plan.lp[k + 1] - plan.lp[k];
          const int ps = g * d.wstride + nu - 1;
          rec[2 * l] = ps | ((nl - 1) << 16);   // candidates: slots ps + 1 .. ps + nl - 1
          rec[2 * l + 1] = k;
          wrows = std::max(wrows, (nu + nl - 1 + TL - 1) / TL);
        }
      }
      // the PIVOT row also does the chunk's write row 0 (its lanes' kinds and positions derived from the PIVOT record:
      // checked here against the WRITE record the row replaces)
      for (int l = 0; l < 64; l++) {
        const int g = l / TL, t = l % TL;
        int want_wk = 0, want_k = 0;
        if (g < ntr) {
          const int k = d.tr_col[t0 + g], nu = plan.up[k + 1] - plan.up[k], nl = plan.lp[k + 1] - plan.lp[k];
          if (t < nu + nl - 1) {
            want_wk = t < nu - 1 ? D::kWriteU : t == nu - 1 ? D::kWritePivot : D::kWriteL;
            want_k = k;
          }
        }
        const D::WriteLane wl = d.pivot_write_lane(rec[2 * l], rec[2 * l + 1], l, TL, n_wr);
        if (wl.wk != want_wk || (want_wk == D::kWritePivot && wl.k != want_k))
          throw std::runtime_error("plan: a PIVOT row's write differs from its WRITE record");
      }
      emit(D::kRsPivot | D::kRsPivotWrite | width, lv, rec.data());
      n_wr++;
      // WRITE rows: (kind << 24 | pivot step for the pivot's lane, vals position); write row w of a chunk writes slot
      // t = w * TL + lane % TL of the lane's track (TL lanes per track) to the stream's next 64 factor positions, an
      // idle lane included (its position is its own and never read); row 0 is the PIVOT row's
      for (int w = 1; w < wrows; w++) {
        for (int l = 0; l < 64; l++) {
          const int g = l / TL, t = w * TL + l % TL;
          rec[2 * l] = 0;
          rec[2 * l + 1] = d.n_a + 64 * n_wr + l;
          if (g >= ntr) continue;
          const int k = d.tr_col[t0 + g], nu = plan.up[k + 1] - plan.up[k];
          const int nl = plan.lp[k + 1] - plan.lp[k];
          if (t >= nu + nl - 1) continue;
          rec[2 * l] = t < nu - 1 ? D::kWriteU << 24 : t == nu - 1 ? (D::kWritePivot << 24) | k : D::kWriteL << 24;
        }
        emit(D::kRsWrite | width, lv, rec.data());
        n_wr++;
      }
      if (d.ch_fence[c]) {
        d.rs_meta.back() |= D::kRsLevelEnd;
        level_end_row[lv] = (int)d.rs_meta.size() - 1;
      }
    }
    if (n_wr != d.n_wrows) throw std::runtime_error("plan: WRITE rows differ from the factor layout");
    d.n_rs = (int)d.rs_meta.size();
    d.rs_wr0.assign(d.n_rs, 0);
    for (int r = 0, c = 0; r < d.n_rs; r++) {
      d.rs_wr0[r] = c;
      const int k = d.rs_meta[r] & D::kRsKindMask;
      c += k == D::kRsWrite || (k == D::kRsPivot && (d.rs_meta[r] & D::kRsPivotWrite));
    }
    d.rs_chunk_row.push_back(d.n_rs);
    while ((int)d.rs_level_chunk.size() <= n_levels) d.rs_level_chunk.push_back(d.n_chunks);
    // fold each row's kind and flags into its records
    for (int r = 0; r < d.n_rs; r++)
      for (int l = 0; l < 64; l++) d.rs_rec[((size_t)r * 64 + l) * 2 + 1] |= d.rs_meta[r] << D::kRsMetaShift;
  }
  // For the exact pivot check: how many candidates (rows not yet pivotal at step k) precede the pivot in reach order.
  d.piv_pos.assign(m, 0);
  for (int k = 0; k < m; k++) {
    int before = 0;
    for (int p = plan.reach_ptr[k]; p < plan.reach_ptr[k + 1]; p++) {
      int row = plan.reach[p];
      if (row == plan.pivot_row[k]) break;
      if (plan.pinv[row] > k) before++;
    }
    d.piv_pos[k] = before;
  }
}

}  // namespace gs

namespace gs {

bool check_level_schedule(const DevicePlanHost& d, int ring, int padded_rows, int padded_fs, int padded_bs,
                          std::string& why) {
  using D = DevicePlanHost;
  auto bad = [&](const std::string& w) { why = w; return false; };
  const int nc = d.n_chunks, nl = d.n_levels;
  if ((int)d.rs_chunk_row.size() != nc + 1 || (int)d.rs_level_chunk.size() != nl + 1) return bad("chunk tables size");
  if (d.rs_chunk_row[0] != 0 || d.rs_chunk_row[nc] != d.n_rs) return bad("chunk rows do not cover the stream");
  if (d.rs_level_chunk[0] != 0 || d.rs_level_chunk[nl] != nc) return bad("level chunks do not cover the chunks");
  const int n_vals = d.n_a + d.n_fac;
  vector<int> val_level(n_vals, -1), val_owner(n_vals, -1), dpos_chunk(d.m, -1);
  for (int l = 0; l < nl; l++) {
    if (d.rs_level_chunk[l + 1] <= d.rs_level_chunk[l]) return bad("a level without chunks");
    // every read of the level first (the level's own writes must not be read by it)
    for (int c = d.rs_level_chunk[l]; c < d.rs_level_chunk[l + 1]; c++) {
      const int r0 = d.rs_chunk_row[c], r1 = d.rs_chunk_row[c + 1];
      if (r1 <= r0) return bad("an empty chunk");
      const int r0e = r0 & ~1;   // the device starts a chunk at the even row of its first row's pair
      const int last_load = r0e + (r1 - r0e + ring - 1) / ring * ring - 1 + ring;
      if (last_load >= padded_rows) return bad("a chunk's look-ahead passes the stream's padding");
      if ((d.rs_meta[r0] & D::kRsKindMask) != D::kRsAssign || !(d.rs_meta[r0] & D::kRsClear))
        return bad("a chunk does not start by clearing its workspace");
      for (int r = r0; r < r1; r++) {
        if (d.rs_level[r] != l) return bad("a chunk row of another level");
        if (r > r0 && (d.rs_meta[r] & D::kRsClear)) return bad("a workspace clear inside a chunk");
        if ((d.rs_meta[r] & D::kRsKindMask) != D::kRsUpdate) continue;
        for (int lane = 0; lane < 64; lane++) {
          const int* rr = &d.rs_rec[((size_t)r * 64 + lane) * 2];
          if ((rr[0] & 0xFFFF) == d.spare(lane)) continue;
          const int v = rr[1] & D::kRsIndexMask;
          if (v < d.n_a || v >= n_vals || val_level[v] < 0 || val_level[v] >= l)
            return bad("a chunk reads an L value its level's predecessors did not write");
        }
      }
    }
    // then the level's writes: disjoint between its chunks
    for (int c = d.rs_level_chunk[l]; c < d.rs_level_chunk[l + 1]; c++) {
      for (int r = d.rs_chunk_row[c]; r < d.rs_chunk_row[c + 1]; r++) {
        const int rk = d.rs_meta[r] & D::kRsKindMask;
        const bool pw = rk == D::kRsPivot && (d.rs_meta[r] & D::kRsPivotWrite);
        if (rk != D::kRsWrite && !pw) continue;
        for (int lane = 0; lane < 64; lane++) {
          const int* rr = &d.rs_rec[((size_t)r * 64 + lane) * 2];
          const D::WriteLane wl = pw ? d.pivot_write_lane(rr[0], rr[1], lane, D::track_lanes(d.rs_meta[r]), d.rs_wr0[r])
                                     : D::WriteLane{(int)((unsigned)rr[0] >> 24), rr[0] & D::kRsIndexMask,
                                                    rr[1] & D::kRsIndexMask};
          const int wk = wl.wk, pos = wl.pos;
          // every lane stores (an idle one too, at a position of its own)
          if (pos < d.n_a || pos >= n_vals) return bad("a factor position outside the factor array");
          if (val_owner[pos] >= 0) return bad("a factor position written twice");
          val_owner[pos] = c;
          if (wk == D::kWriteL) val_level[pos] = l;
          if (wk == D::kWritePivot) {
            int& owner = dpos_chunk.at(wl.k);
            if (owner >= 0) return bad("a pivot written twice");
            owner = c;
          }
        }
      }
    }
  }
  // solve passes: each level's passes cover its steps; a pass reads only rows final before its level
  auto passes = [&](const vector<int>& rec, int steps, const vector<int>& pass_step, const vector<int>& level_pass,
                    int levels, int padded, bool back) -> bool {
    if ((int)level_pass.size() != levels + 1 || pass_step.empty() || pass_step.front() != 0 ||
        pass_step.back() != steps || level_pass[levels] + 1 != (int)pass_step.size())
      return bad("pass tables");
    vector<int> row_level(d.m, -1);
    if (!back)
      for (int i = 0; i < d.m; i++) if (d.lr_ptr[i + 1] == d.lr_ptr[i]) row_level[i] = -2;  // final from the start
    for (int l = 0; l < levels; l++) {
      vector<int> rows;
      for (int p = level_pass[l]; p < level_pass[l + 1]; p++) {
        const int s0 = pass_step[p], s1 = pass_step[p + 1];
        if (s1 <= s0) return bad("an empty pass");
        if (s0 + (s1 - s0 + ring - 1) / ring * ring - 1 + ring >= padded) return bad("a pass's look-ahead passes the padding");
        if ((((unsigned)rec[((size_t)(s1 - 1) * 64) * 2 + 1] >> D::kRsMetaShift) & D::kStepKindMask) != D::kStepChain)
          return bad("a pass does not end on its own CHAIN step");
        for (int t = s0; t < s1; t++)
          for (int lane = 0; lane < 64; lane++) {
            const int* r2 = &rec[((size_t)t * 64 + lane) * 2];
            const int meta = (int)((unsigned)r2[1] >> D::kRsMetaShift), kind = meta & D::kStepKindMask;
            const int field = r2[1] & D::kRsIndexMask;
            if (kind == D::kStepProduct && r2[0] >= 0 && (row_level.at(r2[0]) == -1 || row_level[r2[0]] >= l))
              return bad("a pass reads a row its level's predecessors did not finish");
            if (kind == D::kStepChain && !(meta & D::kChainIdle) && (meta & D::kChainEnd)) rows.push_back(field);
          }
      }
      for (int r : rows) {
        if (row_level.at(r) >= 0) return bad("a row solved twice");
        row_level[r] = l;
      }
    }
    return true;
  };
  if (!passes(d.fs_rec, d.fs_steps, d.fs_pass_step, d.fs_level_pass, d.fw_levels, padded_fs, false)) return false;
  if (!passes(d.bs_rec, d.bs_steps, d.bs_pass_step, d.bs_level_pass, d.bw_levels, padded_bs, true)) return false;
  return true;
}

bool emulate_device_lu(const DevicePlanHost& d, int nnz_j, const double* jv, const double* rhs, double* out,
                       std::string& why) {
  const int m = d.m, G = DevicePlanHost::kTracks;
  const int wsz = G * d.wstride;
  vector<double> ws(wsz), dval(m, std::nan(""));
  auto bad = [&](const std::string& w) { why = w; return false; };
  using D = DevicePlanHost;
  // vals = [A values at their positions, as the mismatch kernel writes them (j2a) | the factors, 64 per WRITE row]
  const int n_a = d.n_a, n_vals = n_a + d.n_fac;
  vector<double> vals(n_vals, std::nan(""));
  for (int s = 0; s < nnz_j && s < (int)d.j2a.size(); s++) {
    int e = d.j2a[s];
    if (e < -1 || e >= n_a) return bad("j2a index");
    if (e >= 0) vals[e] = jv[s];
  }
  if ((int)d.rs_meta.size() != d.n_rs || (int)d.rs_rec.size() != 2 * 64 * d.n_rs) return bad("row stream size");
  // The device loads row r's value at the end of row r - kRsValueAhead (or before the first row). fence_row[v] is the
  // row whose level-end fence publishes vals[v] (-1: final before the refactor starts).
  vector<int> fence_row(n_vals, -1), written_row(n_vals, -1);
  vector<int> pending;  // values written in the current level, published at its level end
  vector<double> piv(64, 0.0);
  int write_row = 0;   // WRITE rows since the chunk's PIVOT row
  // the reference's update order: per chunk, each workspace slot's updates come in increasing U entry t, and a slot
  // is read (as an update's source) only after all of its own updates
  vector<int> n_pending(wsz, 0), last_t(wsz, -1);
  for (int r = 0; r < d.n_rs; r++) {
    const int meta = d.rs_meta[r], kind = meta & D::kRsKindMask;
    if (kind == D::kRsAssign && (meta & D::kRsClear)) {
      std::fill(ws.begin(), ws.end(), 0.0);
      std::fill(n_pending.begin(), n_pending.end(), 0);
      std::fill(last_t.begin(), last_t.end(), -1);
      for (int r2 = r; r2 < d.n_rs; r2++) {
        if (r2 > r && (d.rs_meta[r2] & D::kRsKindMask) == D::kRsAssign && (d.rs_meta[r2] & D::kRsClear)) break;
        if ((d.rs_meta[r2] & D::kRsKindMask) != D::kRsUpdate) continue;
        for (int lane = 0; lane < 64; lane++) {
          const int x = d.rs_rec[((size_t)r2 * 64 + lane) * 2], dst = x & 0xFFFF;
          if (dst < wsz) n_pending[dst]++;
        }
      }
    }
    if (kind == D::kRsUpdate) {
      vector<char> seen(wsz, 0);
      for (int lane = 0; lane < 64; lane++) {
        const int x = d.rs_rec[((size_t)r * 64 + lane) * 2], dst = x & 0xFFFF, src = (int)((unsigned)x >> 16);
        if (dst == d.spare(lane)) continue;
        if (dst >= wsz || src >= wsz) return bad("row record outside the workspace");
        const int t = src % d.wstride;
        if (seen[dst]) return bad("two lanes of a row update one slot");
        seen[dst] = 1;
        if (n_pending[src] != 0) return bad("an update reads a slot before all of the slot's updates");
        if (t <= last_t[dst]) return bad("a slot's updates out of the reference's order");
      }
      for (int lane = 0; lane < 64; lane++) {
        const int x = d.rs_rec[((size_t)r * 64 + lane) * 2], dst = x & 0xFFFF, src = (int)((unsigned)x >> 16);
        if (dst == d.spare(lane)) continue;
        n_pending[dst]--;
        last_t[dst] = src % d.wstride;
      }
    }
    bool fail = false;
    const int TL = D::track_lanes(meta);   // lanes per track of the row's chunk
    if (kind == D::kRsPivot) write_row = 0;
    // a lane's WRITE: its slot of write row write_row to vals[pos] (an idle lane's store is never read: NaN here)
    std::string werr;
    auto do_write = [&](int lane, int wk, int k, int pos) -> bool {
      const int slot = (lane / TL) * d.wstride + write_row * TL + lane % TL;
      if (slot >= wsz + D::kSpareSlots) { werr = "write slot outside the workspace"; return false; }
      if (pos < n_a || pos >= n_vals) { werr = "write position"; return false; }
      if (wk == 0) {
        vals[pos] = std::nan("");
        return true;
      }
      if (slot >= wsz || slot / d.wstride != lane / TL) { werr = "write slot"; return false; }
      if (wk == D::kWriteU) {
        vals[pos] = ws[slot];
      } else if (wk == D::kWritePivot) {
        if (k >= m) { werr = "pivot write position"; return false; }
        vals[pos] = dval[k] = ws[slot];
      } else if (wk == D::kWriteL) {
        if (piv[lane] == 0.0) { werr = "L write without a pivot"; return false; }
        vals[pos] = ws[slot] / piv[lane];
        written_row[pos] = r;
        pending.push_back(pos);
      } else {
        werr = "write kind";
        return false;
      }
      return true;
    };
    for (int lane = 0; lane < 64; lane++) {
      const int* rr = &d.rs_rec[((size_t)r * 64 + lane) * 2];
      if ((int)((unsigned)rr[1] >> D::kRsMetaShift) != meta) return bad("row kind differs across lanes");
      const int rc[2] = {rr[0], rr[1] & D::kRsIndexMask};
      if ((kind == D::kRsAssign || kind == D::kRsUpdate) && rc[1] >= n_vals) return bad("row value index");
      // the device loads vals[field] for every row (rows without a value too): vals holds max(n_vals, m)
      if (rc[1] >= std::max(n_vals, m)) return bad("row field outside the value array");
      const int load_row = r - D::kRsValueAhead;  // the value is loaded at the end of this row
      if (kind == D::kRsAssign) {
        if (rc[0] == d.spare(lane)) continue;
        if (rc[0] < 0 || rc[0] >= wsz) return bad("A workspace index");  // any lane may assign any slot
        if (rc[1] >= n_a || std::isnan(vals[rc[1]])) return bad("A value not written by the Jacobian fill");
        ws[rc[0]] = vals[rc[1]];
      } else if (kind == D::kRsUpdate) {
        const int dst = rc[0] & 0xFFFF, xj = rc[0] >> 16;
        if (dst == d.spare(lane) && xj == d.spare(lane)) continue;
        if (dst >= wsz || xj >= wsz || dst / d.wstride != lane / TL || xj / d.wstride != lane / TL)
          return bad("row record outside its track's workspace");
        if (rc[1] < n_a || written_row[rc[1]] < 0) return bad("row reads an L value not yet written");
        if (fence_row[rc[1]] < 0 || fence_row[rc[1]] > load_row) return bad("row value loaded before its fence");
        ws[dst] = ws[dst] - vals[rc[1]] * ws[xj];
      } else if (kind == D::kRsPivot) {
        if (rc[0] != -1) {
          const int ps = rc[0] & 0xFFFF, lo = ps + 1, hi = ps + 1 + (int)((unsigned)rc[0] >> 16);
          if (ps < 0 || ps >= wsz || ps / d.wstride != lane / TL || lo != ps + 1 || hi < lo ||
              hi > (lane / TL + 1) * d.wstride)
            return bad("pivot record");
          piv[lane] = ws[ps];
          fail |= piv[lane] == 0.0;
        }
        if (meta & D::kRsPivotWrite) {
          const D::WriteLane wl = d.pivot_write_lane(rr[0], rr[1], lane, TL, d.rs_wr0[r]);
          if (!do_write(lane, wl.wk, wl.k, wl.pos)) return bad(werr);
        }
      } else if (kind == D::kRsWrite) {
        if (rc[1] != n_a + 64 * d.rs_wr0[r] + lane) return bad("write position out of the stream's order");
        if (!do_write(lane, (int)((unsigned)rc[0] >> 24), rc[0] & D::kRsIndexMask, rc[1])) return bad(werr);
      } else if (kind != D::kRsNop) {
        return bad("row kind");
      }
    }
    if (fail) return bad("zero pivot");
    if (kind == D::kRsWrite || (kind == D::kRsPivot && (meta & D::kRsPivotWrite))) write_row++;
    if (meta & D::kRsLevelEnd) {
      for (int v : pending) fence_row[v] = r;
      pending.clear();
    }
  }
  if (!pending.empty()) return bad("values written after the last level end");
  vector<double> x(m, std::nan(""));
  for (int r = 0; r < m; r++) x[d.pinv[r]] = rhs[r];
  // the step streams, with the rows a step may read checked against the rows already final
  auto run_stream = [&](const vector<int>& rec, int steps, const vector<double>& vals, bool back,
                        vector<char>& final_rows) -> bool {
    using D = DevicePlanHost;
    const int cap = d.solve_scratch();
    vector<double> acc(64, 0.0), scratch(cap, std::nan(""));
    vector<int> row(64, -1), end_step(m, -D::kStepXAhead);
    int sbase = 0;
    if ((int)rec.size() != steps * 64 * 2) { why = "stream size"; return false; }
    for (int t = 0; t < steps; t++) {
      const int kind = (int)((unsigned)rec[((size_t)t * 64) * 2 + 1] >> D::kRsMetaShift) & D::kStepKindMask;
      vector<int> ended;
      for (int lane = 0; lane < 64; lane++) {
        const int* r2 = &rec[((size_t)t * 64 + lane) * 2];
        const int meta = (int)((unsigned)r2[1] >> D::kRsMetaShift), field = r2[1] & D::kRsIndexMask;
        if ((meta & D::kStepKindMask) != kind) { why = "stream step kind differs across lanes"; return false; }
        if (kind == D::kStepProduct && meta != (int)((unsigned)rec[((size_t)t * 64) * 2 + 1] >> D::kRsMetaShift))
          { why = "product flags differ across lanes"; return false; }
        // the device loads the value array at every step's field (it holds max(values, m)); the pivot array at a
        // CHAIN step's row only
        if (field >= std::max({(int)vals.size(), m, 1})) { why = "step field outside the value array"; return false; }
        if (kind == D::kStepProduct) {
          if (sbase + 64 > cap) { why = "products pass the scratch area"; return false; }
          if (r2[0] < 0) continue;
          if (r2[0] >= m || !final_rows[r2[0]]) { why = "stream reads a row not yet final"; return false; }
          // a device with x in global memory loads it kStepXAhead steps early unless the step says otherwise
          if (!(meta & D::kProductLate) && end_step[r2[0]] > t - D::kStepXAhead)
            { why = "a product loads x ahead of its column's last chain"; return false; }
          if (field >= (int)vals.size()) { why = "stream value index"; return false; }
          scratch[sbase + lane] = vals[field] * x[r2[0]];
        } else if (kind == D::kStepChain) {
          if (field >= m) { why = "chain row"; return false; }
          if (meta & D::kChainIdle) {
            if (meta & (D::kChainBegin | D::kChainEnd)) { why = "an idle chain lane begins or ends a row"; return false; }
            continue;
          }
          const int st = r2[0] & 0xFFFF, len = (int)((unsigned)r2[0] >> 16);
          if (st + len > cap) { why = "chain outside the scratch area"; return false; }
          if (meta & D::kChainBegin) {
            row[lane] = field;
            acc[lane] = x[row[lane]];
          } else if (row[lane] != field) {
            why = "a chain continues another lane's row";
            return false;
          }
          for (int e = st; e < st + len; e++) {
            if (std::isnan(scratch[e])) { why = "chain reads a product not formed"; return false; }
            acc[lane] = acc[lane] - scratch[e];
          }
          if (meta & D::kChainEnd) {
            x[row[lane]] = back ? acc[lane] / dval[row[lane]] : acc[lane];
            ended.push_back(row[lane]);
            end_step[row[lane]] = t;
            row[lane] = -1;
          }
        } else if (kind != 0) {
          why = "stream step kind";
          return false;
        }
      }
      if (kind == D::kStepProduct) sbase += 64;
      if (kind == D::kStepChain) {
        sbase = 0;
        std::fill(scratch.begin(), scratch.end(), std::nan(""));
        for (int r : ended) final_rows[r] = 1;
      }
    }
    return true;
  };
  {
    // forward: rows without L entries are final from the start; back: a row is final once divided
    vector<char> fin(m, 0);
    for (int i = 0; i < m; i++) fin[i] = d.lr_ptr[i + 1] == d.lr_ptr[i];
    const vector<double> fac(vals.begin() + n_a, vals.end());
    if (!run_stream(d.fs_rec, d.fs_steps, fac, false, fin)) return false;
    vector<char> fin_b(m, 0);
    if (!run_stream(d.bs_rec, d.bs_steps, fac, true, fin_b)) return false;
    for (int k = 0; k < m; k++) out[d.q[k]] = x[k];
    return true;
  }
}

}  // namespace gs
