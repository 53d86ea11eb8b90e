// mechgen.cpp -- see mechgen.h.
//
// Arithmetic conventions of the emitted code, kept identical to the oracle so
// that the natural-order plan reproduces the reference's dense pivot-free
// elimination (rx_traced_integrator.jl:110-135) operation for operation:
//   rate_r   = (k_r * u_a) * u_b, k_r -> k_r * jfac for photolysis;
//   f_i      = e_i, then f_i += nu_ir * rate_r in reaction order;
//   J_ia     = 0, then J_ia += nu_ir * d rate_r / d u_a in reaction order;
//   W_ia     = (i == a ? 1.0 : 0.0) - (h gamma) * J_ia   (:291-296);
//   LU       m = W_ik * (1 / W_kk); W_ij = W_ij - m * W_kj   (:113-124);
//   solve    forward y_i = b_i - sum_k m_ik y_k (k ascending), backward
//            x_i = (y_i - sum_j U_ij x_j) / U_ii (j ascending)   (:126-133).
// Elimination skips only structural zeros, whose products are exact zeros.
#include "mechgen.h"

#include <algorithm>
#include <sstream>
#include <stdexcept>

namespace rs {

std::string validate_mech(const MechSpec& m) {
    if (m.ns < 1) return "mechanism needs at least one species";
    if (m.nr < 0) return "negative reaction count";
    auto bad = [&](int x, bool allow_none) { return x < (allow_none ? -1 : 0) || x >= m.ns; };
    if (!m.prod.empty() && (int)m.prod.size() != m.nr) return "product lists: one per reaction";
    for (int r = 0; r < m.nr; ++r) {
        if (bad(m.ra[r], false) || bad(m.rb[r], true))
            return "species index out of range in reaction " + std::to_string(r);
        if (m.prod.empty()) {
            if (bad(m.pa[r], true) || bad(m.pb[r], true))
                return "species index out of range in reaction " + std::to_string(r);
        } else {
            for (auto& q : m.prod[r]) {
                if (bad(q.first, false))
                    return "product species out of range in reaction " + std::to_string(r);
                if (!(q.second > 0.0) || !(q.second < 1e6))
                    return "product coefficient must be positive and finite in reaction " + std::to_string(r);
            }
        }
        if (m.photo[r] && m.rb[r] >= 0)
            return "photolysis reaction " + std::to_string(r) + " must be first order";
    }
    return "";
}

// Net stoichiometry per reaction, species ascending: substrates count -1 each
// (a species twice for a second-order self-reaction), products their
// coefficient; zero entries (a catalyst) are dropped.
static std::vector<std::vector<std::pair<int, double>>> net_stoich(const MechSpec& m) {
    std::vector<std::vector<std::pair<int, double>>> net(m.nr);
    for (int r = 0; r < m.nr; ++r) {
        std::vector<double> nu(m.ns, 0.0);
        nu[m.ra[r]] -= 1.0;
        if (m.rb[r] >= 0) nu[m.rb[r]] -= 1.0;
        if (m.prod.empty()) {
            if (m.pa[r] >= 0) nu[m.pa[r]] += 1.0;
            if (m.pb[r] >= 0) nu[m.pb[r]] += 1.0;
        } else {
            for (auto& q : m.prod[r]) nu[q.first] += q.second;
        }
        for (int s = 0; s < m.ns; ++s)
            if (nu[s] != 0.0) net[r].emplace_back(s, nu[s]);
    }
    return net;
}

MechPlan plan_mechanism(const MechSpec& m, int ordering) {
    MechPlan p;
    p.ns = m.ns;
    p.nr = m.nr;
    p.ordering = ordering;
    p.net = net_stoich(m);
    const int n = m.ns;
    // Jacobian pattern (species coordinates)
    p.jpat.assign(n, std::vector<char>(n, 0));
    for (int r = 0; r < m.nr; ++r) {
        for (auto& pr : p.net[r]) {
            p.jpat[pr.first][m.ra[r]] = 1;
            if (m.rb[r] >= 0) p.jpat[pr.first][m.rb[r]] = 1;
        }
    }
    // W pattern = J pattern + diagonal
    std::vector<std::vector<char>> A(n, std::vector<char>(n, 0));
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < n; ++a) A[i][a] = p.jpat[i][a] || i == a;
    // elimination order
    p.perm.clear();
    if (ordering == 0) {
        for (int i = 0; i < n; ++i) p.perm.push_back(i);
    } else {
        // greedy Markowitz: minimise (row count - 1) * (col count - 1) over the
        // remaining active submatrix, ties to the lowest species index.
        std::vector<std::vector<char>> B = A;
        std::vector<char> done(n, 0);
        for (int step = 0; step < n; ++step) {
            long long best = -1;
            int bv = -1;
            for (int v = 0; v < n; ++v) {
                if (done[v]) continue;
                long long rc = 0, cc = 0;
                for (int x = 0; x < n; ++x) {
                    if (done[x] || x == v) continue;
                    rc += B[v][x];
                    cc += B[x][v];
                }
                long long score = rc * cc;
                if (bv < 0 || score < best) { best = score; bv = v; }
            }
            for (int i = 0; i < n; ++i) {
                if (done[i] || i == bv || !B[i][bv]) continue;
                for (int j = 0; j < n; ++j)
                    if (!done[j] && j != bv && B[bv][j]) B[i][j] = 1;
            }
            done[bv] = 1;
            p.perm.push_back(bv);
        }
    }
    p.pos.assign(n, 0);
    for (int q = 0; q < n; ++q) p.pos[p.perm[q]] = q;
    // symbolic LU in position coordinates
    std::vector<std::vector<char>> P(n, std::vector<char>(n, 0));
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) P[i][j] = A[p.perm[i]][p.perm[j]];
    for (int k = 0; k < n; ++k)
        for (int i = k + 1; i < n; ++i) {
            if (!P[i][k]) continue;
            for (int j = k + 1; j < n; ++j)
                if (P[k][j]) P[i][j] = 1;
        }
    p.idx.assign(n, std::vector<int>(n, -1));
    p.nnz = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if (P[i][j]) p.idx[i][j] = p.nnz++;
    // work counts
    MechWork& w = p.work;
    w.nnz_LU = p.nnz;
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < n; ++a) w.nnz_J += p.jpat[i][a];
    for (int r = 0; r < m.nr; ++r) {
        const int nre = m.rb[r] >= 0 ? 2 : 1;
        const long long nn = (long long)p.net[r].size();
        w.f_rhs += (nre == 2 ? 2 : 1) + (m.photo[r] ? 1 : 0) + 2 * nn;
        const bool same = m.rb[r] >= 0 && m.rb[r] == m.ra[r];
        const int nder = (nre == 2 && !same) ? 2 : 1;
        w.f_jac += (nre == 2 ? (same ? 2 : 2) : 0) + (m.photo[r] ? 1 : 0) + 2 * nn * nder;
        w.f_jtv += 2 * nn + (nre == 2 ? (same ? 2 : 4) : 0) + 2 * nder + (m.photo[r] ? 1 : 0);
        if (nre == 2) w.f_hess += 2 * nn + (same ? 4 : 8);
    }
    w.f_w = 2 * w.nnz_J;
    for (int k = 0; k < n; ++k) {
        w.f_lu += 1;
        long long ncols = 0;
        for (int j = k + 1; j < n; ++j) ncols += (p.idx[k][j] >= 0);
        for (int i = k + 1; i < n; ++i)
            if (p.idx[i][k] >= 0) w.f_lu += 1 + 2 * ncols;
    }
    long long offd = p.nnz - n;
    w.f_solve = 2 * offd + n;
    w.f_solveT = 2 * offd + n;
    // per-species vector work: u1 (2), b2 (1), k2 (1), unew (2), b3 (6), error (9)
    w.f_vec_attempt = 21LL * n;
    // vjp: u1 (2), b2 (1), k2 (1), kbar2 (1), kbar1 (1), ubar/k1bar (3), v (1), e (2)
    w.f_vec_vjp = 12LL * n;
    return p;
}

namespace {

struct Emit {
    std::ostringstream o;
    template <class T>
    Emit& operator<<(const T& x) { o << x; return *this; }
};

std::string kexpr(const MechSpec& m, int r) {
    std::ostringstream s;
    if (m.photo[r]) s << "(K[" << r << "] * jf)";
    else s << "K[" << r << "]";
    return s.str();
}

// "x = x + nu*y" with the exact forms of the oracle's `x += (double)nu * y`.
// x = x + nu * y with the coefficient as an exact literal (x - y for -1).
std::string accum(const std::string& x, double nu, const std::string& y) {
    std::ostringstream s;
    s.precision(17);
    if (nu == 1.0) s << x << " = " << x << " + " << y << ";";
    else if (nu == -1.0) s << x << " = " << x << " - " << y << ";";
    else if (nu > 0) s << x << " = " << x << " + " << std::showpoint << nu << " * " << y << ";";
    else s << x << " = " << x << " - " << std::showpoint << (-nu) << " * " << y << ";";
    return s.str();
}

}  // namespace

std::string generate_mech_source(const MechSpec& m, const MechPlan& p) {
    const int n = p.ns;
    Emit e;
    e << "// generated by mechgen: " << n << " species, " << m.nr << " reactions, "
      << "W/LU non-zeros " << p.nnz << ", ordering "
      << (p.ordering ? "Markowitz" : "natural") << "\n";
    e << "#define RS_NS " << n << "\n#define RS_NR " << m.nr << "\n#define RS_NNZ " << p.nnz
      << "\n";
    // ---- RHS: rates folded into the accumulation, reaction order per species
    e << "template <int US = 1>\n";
    e << "RS_FN void rs_rhs(const double* __restrict__ K, double jf, const double* __restrict__ E,"
         " const double* __restrict__ u, double* __restrict__ f) {\n";
    // emissions of this cell (E: its NS values, contiguous), one branch for all
    e << "  if (E) {\n";
    for (int i = 0; i < n; ++i) e << "    f[" << i << "] = E[" << i << "];\n";
    e << "  } else {\n";
    for (int i = 0; i < n; ++i) e << "    f[" << i << "] = 0.0;\n";
    e << "  }\n";
    for (int r = 0; r < m.nr; ++r) {
        e << "  { const double rt = ";
        if (m.rb[r] < 0) e << kexpr(m, r) << " * u[" << m.ra[r] << " * US];";
        else e << "(" << kexpr(m, r) << " * u[" << m.ra[r] << " * US]) * u[" << m.rb[r] << " * US];";
        for (auto& pr : p.net[r]) e << " " << accum("f[" + std::to_string(pr.first) + "]", pr.second, "rt");
        e << " }\n";
    }
    e << "}\n";
    // ---- W = I - hg*J directly in LU slot storage: J is accumulated in the
    // slots of its pattern (each entry in reaction order), then turned into W
    e << "template <int US = 1, int WS = 1>\n";
    e << "RS_FN void rs_jac_W(const double* __restrict__ K, double jf, const double* __restrict__ u,"
         " double hg, double* __restrict__ W) {\n";
    std::vector<std::vector<int>> slot(n, std::vector<int>(n, -1));   // species (i, a) -> LU slot
    for (int qi = 0; qi < n; ++qi)
        for (int qj = 0; qj < n; ++qj)
            if (p.idx[qi][qj] >= 0) slot[p.perm[qi]][p.perm[qj]] = p.idx[qi][qj];
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < n; ++a)
            if (p.jpat[i][a] && slot[i][a] < 0)
                throw std::logic_error("Jacobian entry outside the LU pattern");
    auto wname = [&](int i, int a) { return "W[" + std::to_string(slot[i][a]) + " * WS]"; };
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < n; ++a)
            if (p.jpat[i][a]) e << "  " << wname(i, a) << " = 0.0;\n";
    for (int r = 0; r < m.nr; ++r) {
        int a = m.ra[r], b = m.rb[r];
        e << "  { const double kc = " << kexpr(m, r) << ";";
        if (b < 0) e << " const double da = kc;";
        else if (a == b) e << " const double da = (2.0 * kc) * u[" << a << " * US];";
        else e << " const double da = kc * u[" << b << " * US]; const double db = kc * u[" << a << " * US];";
        for (auto& pr : p.net[r]) {
            e << " " << accum(wname(pr.first, a), pr.second, "da");
            if (b >= 0 && b != a) e << " " << accum(wname(pr.first, b), pr.second, "db");
        }
        e << " }\n";
    }
    for (int qi = 0; qi < n; ++qi)
        for (int qj = 0; qj < n; ++qj) {
            int sl = p.idx[qi][qj];
            if (sl < 0) continue;
            int i = p.perm[qi], a = p.perm[qj];
            const std::string w = "W[" + std::to_string(sl) + " * WS]";
            if (p.jpat[i][a])
                e << "  " << w << " = " << (i == a ? "1.0" : "0.0") << " - hg * " << w << ";\n";
            else
                e << "  " << w << " = " << (i == a ? "1.0" : "0.0") << ";\n";
        }
    e << "}\n";
    // ---- LU without pivoting, in position order
    e << "template <int WS = 1>\n";
    e << "RS_FN void rs_lu(double* __restrict__ W) {\n";
    for (int k = 0; k < n; ++k) {
        e << "  { const double pv = RS_PIVOT(W[" << p.idx[k][k] << " * WS]);\n";
        for (int i = k + 1; i < n; ++i) {
            int sik = p.idx[i][k];
            if (sik < 0) continue;
            e << "  { const double m = W[" << sik << " * WS] * pv; W[" << sik << " * WS] = m;";
            for (int j = k + 1; j < n; ++j) {
                int skj = p.idx[k][j];
                if (skj < 0) continue;
                int sij = p.idx[i][j];
                e << " W[" << sij << " * WS] = W[" << sij << " * WS] - m * W[" << skj << " * WS];";
            }
            e << " }\n";
        }
        e << "  }\n";
    }
    e << "}\n";
    // ---- W x = b  (b, x in species coordinates)
    e << "template <int WS = 1>\n";
    e << "RS_FN void rs_solve(const double* __restrict__ W,"
         " const double* __restrict__ b, double* __restrict__ x) {\n";
    for (int q = 0; q < n; ++q) {
        e << "  double y" << q << " = b[" << p.perm[q] << "];";
        for (int k = 0; k < q; ++k)
            if (p.idx[q][k] >= 0) e << " y" << q << " = y" << q << " - W[" << p.idx[q][k] << " * WS] * y" << k << ";";
        e << "\n";
    }
    for (int q = n - 1; q >= 0; --q) {
        e << "  { double acc = y" << q << ";";
        for (int j = q + 1; j < n; ++j)
            if (p.idx[q][j] >= 0) e << " acc = acc - W[" << p.idx[q][j] << " * WS] * x[" << p.perm[j] << "];";
        e << " x[" << p.perm[q] << "] = RS_DIAG_SOLVE(acc, W[" << p.idx[q][q] << " * WS]); }\n";
    }
    e << "}\n";
    // ---- W x = b in place (b becomes x; the same operations as rs_solve)
    e << "template <int WS = 1>\n";
    e << "RS_FN void rs_solve_ip(const double* __restrict__ W, double* __restrict__ b) {\n";
    for (int q = 0; q < n; ++q) {
        bool any = false;
        for (int k = 0; k < q; ++k) any = any || p.idx[q][k] >= 0;
        if (!any) continue;
        e << "  { double y = b[" << p.perm[q] << "];";
        for (int k = 0; k < q; ++k)
            if (p.idx[q][k] >= 0) e << " y = y - W[" << p.idx[q][k] << " * WS] * b[" << p.perm[k] << "];";
        e << " b[" << p.perm[q] << "] = y; }\n";
    }
    for (int q = n - 1; q >= 0; --q) {
        e << "  { double acc = b[" << p.perm[q] << "];";
        for (int j = q + 1; j < n; ++j)
            if (p.idx[q][j] >= 0) e << " acc = acc - W[" << p.idx[q][j] << " * WS] * b[" << p.perm[j] << "];";
        e << " b[" << p.perm[q] << "] = RS_DIAG_SOLVE(acc, W[" << p.idx[q][q] << " * WS]); }\n";
    }
    e << "}\n";
    // ---- W^T x = b: U^T z = b (forward), then L^T x = z (backward)
    e << "template <int WS = 1>\n";
    e << "RS_FN void rs_solveT(const double* __restrict__ W,"
         " const double* __restrict__ b, double* __restrict__ x) {\n";
    for (int q = 0; q < n; ++q) {
        e << "  double z" << q << "; { double acc = b[" << p.perm[q] << "];";
        for (int k = 0; k < q; ++k)
            if (p.idx[k][q] >= 0) e << " acc = acc - W[" << p.idx[k][q] << " * WS] * z" << k << ";";
        e << " z" << q << " = RS_DIAG_SOLVE(acc, W[" << p.idx[q][q] << " * WS]); }\n";
    }
    for (int q = n - 1; q >= 0; --q) {
        e << "  { double acc = z" << q << ";";
        for (int i = q + 1; i < n; ++i)
            if (p.idx[i][q] >= 0) e << " acc = acc - W[" << p.idx[i][q] << " * WS] * x[" << p.perm[i] << "];";
        e << " x[" << p.perm[q] << "] = acc; }\n";
    }
    e << "}\n";
    // ---- out += J(u)^T z
    e << "template <int US = 1>\n";
    e << "RS_FN void rs_jtv(const double* __restrict__ K, double jf, const double* __restrict__ u,"
         " const double* __restrict__ z, double* __restrict__ out) {\n";
    for (int r = 0; r < m.nr; ++r) {
        if (p.net[r].empty()) continue;
        int a = m.ra[r], b = m.rb[r];
        e << "  { double s = 0.0;";
        for (auto& pr : p.net[r]) e << " " << accum("s", pr.second, "z[" + std::to_string(pr.first) + "]");
        e << " const double kc = " << kexpr(m, r) << ";";
        if (b < 0) e << " out[" << a << "] = out[" << a << "] + kc * s;";
        else if (a == b) e << " out[" << a << "] = out[" << a << "] + ((2.0 * kc) * u[" << a << " * US]) * s;";
        else
            e << " out[" << a << "] = out[" << a << "] + (kc * u[" << b << " * US]) * s; out[" << b
              << "] = out[" << b << "] + (kc * u[" << a << " * US]) * s;";
        e << " }\n";
    }
    e << "}\n";
    // ---- o1 += J(u)^T z and o2 += c2 * J(u)^T z in one pass (adjoint of u1)
    e << "template <int US = 1>\n";
    e << "RS_FN void rs_jtv2(const double* __restrict__ K, double jf, const double* __restrict__ u,"
         " const double* __restrict__ z, double* __restrict__ o1, double* __restrict__ o2,"
         " double c2) {\n";
    for (int r = 0; r < m.nr; ++r) {
        if (p.net[r].empty()) continue;
        int a = m.ra[r], b = m.rb[r];
        e << "  { double s = 0.0;";
        for (auto& pr : p.net[r]) e << " " << accum("s", pr.second, "z[" + std::to_string(pr.first) + "]");
        e << " const double kc = " << kexpr(m, r) << ";";
        auto both = [&](int x, const std::string& d) {
            e << " { const double t = " << d << " * s; o1[" << x << "] = o1[" << x << "] + t; o2[" << x
              << "] = o2[" << x << "] + c2 * t; }";
        };
        if (b < 0) both(a, "kc");
        else if (a == b) both(a, "((2.0 * kc) * u[" + std::to_string(a) + " * US])");
        else {
            both(a, "(kc * u[" + std::to_string(b) + " * US])");
            both(b, "(kc * u[" + std::to_string(a) + " * US])");
        }
        e << " }\n";
    }
    e << "}\n";
    // ---- ub += coef * d/du [ z^T J(u) v ]   (bimolecular reactions only)
    e << "template <int VS = 1>\n";
    e << "RS_FN void rs_hess(const double* __restrict__ K, double jf, const double* __restrict__ z,"
         " const double* __restrict__ v, double coef, double* __restrict__ ub) {\n";
    for (int r = 0; r < m.nr; ++r) {
        int a = m.ra[r], b = m.rb[r];
        if (b < 0 || p.net[r].empty()) continue;
        e << "  { double s = 0.0;";
        for (auto& pr : p.net[r]) e << " " << accum("s", pr.second, "z[" + std::to_string(pr.first) + "]");
        e << " const double cs = (coef * " << kexpr(m, r) << ") * s;";
        if (a == b) e << " ub[" << a << "] = ub[" << a << "] + (2.0 * cs) * v[" << a << " * VS];";
        else
            e << " ub[" << a << "] = ub[" << a << "] + cs * v[" << b << " * VS]; ub[" << b << "] = ub["
              << b << "] + cs * v[" << a << " * VS];";
        e << " }\n";
    }
    e << "}\n";
    return e.o.str();
}

}  // namespace rs
