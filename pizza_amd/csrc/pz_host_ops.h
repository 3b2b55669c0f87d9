// Host-side set-up of libpizza_amd: sparse-Chebyshev operator stencils,
// radial grid and radial functions.  Included once by pz_api.hip.
//
// Stencils: row idx of I^J [ sum_k p_k(r) D^k ] in the Chebyshev basis with
// halved end coefficients, evaluated by composing the tridiagonal
// multiplication-by-r operator (r = a x + b) and the integration operator,
// and reducing each D^k with I(p D g) = p g - I(p' g).  This reproduces the
// closed-form stencils of chebsparselib.f90 (Marti et al. 2016); negative
// columns are folded onto positive ones like mirror_stencil
// (chebsparselib.f90:997-1020).  Evaluated in long double, rounded once.
#pragma once
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

namespace pzh {

typedef long double ld;

// Laurent polynomial in r: exponent e -> coefficient
struct Laurent {
    std::map<int, ld> c;
    static Laurent mono(ld v, int e) { Laurent l; l.c[e] = v; return l; }
    Laurent operator+(const Laurent &o) const { Laurent r = *this; for (auto &kv : o.c) r.c[kv.first] += kv.second; return r; }
    Laurent operator*(ld s) const { Laurent r; for (auto &kv : c) r.c[kv.first] = kv.second * s; return r; }
    Laurent operator*(const Laurent &o) const
    {
        Laurent r;
        for (auto &a : c)
            for (auto &b : o.c) r.c[a.first + b.first] += a.second * b.second;
        return r;
    }
    Laurent deriv() const
    {
        Laurent r;
        for (auto &kv : c)
            if (kv.first != 0 && kv.second != 0) r.c[kv.first - 1] += kv.second * (ld)kv.first;
        return r;
    }
    bool zero() const { for (auto &kv : c) if (kv.second != 0) return false; return true; }
};

// differential operator sum_k a[k](r) D^k
struct DiffOp {
    std::vector<Laurent> a;
    DiffOp() : a(8) {}
    DiffOp operator+(const DiffOp &o) const { DiffOp r; for (int k = 0; k < 8; k++) r.a[k] = a[k] + o.a[k]; return r; }
    DiffOp scaled(const Laurent &p) const { DiffOp r; for (int k = 0; k < 8; k++) r.a[k] = p * a[k]; return r; }
    // (this o B)
    DiffOp compose(const DiffOp &B) const
    {
        DiffOp R;
        for (int k = 0; k < 8; k++) {
            if (a[k].zero()) continue;
            for (int l = 0; l < 8; l++) {
                if (B.a[l].zero()) continue;
                Laurent bl = B.a[l];
                ld binom = 1;
                for (int i = 0; i <= k; i++) {
                    if (i > 0) { binom = binom * (ld)(k - i + 1) / (ld)i; bl = bl.deriv(); }
                    R.a[k - i + l] = R.a[k - i + l] + (a[k] * bl) * binom;
                }
            }
        }
        return R;
    }
};

// sequences on the index range [-E, L)
struct SeqOps {
    int E, L, T;
    ld a, b;
    void mulr(const std::vector<ld> &y, std::vector<ld> &o) const
    {
        for (int i = 0; i < T; i++) {
            ld ym = i > 0 ? y[i - 1] : 0, yp = i + 1 < T ? y[i + 1] : 0;
            o[i] = a * (ym + yp) / 2 + b * y[i];
        }
    }
    void integ(const std::vector<ld> &y, std::vector<ld> &o) const
    {
        for (int i = 0; i < T; i++) {
            int n = i - E;
            if (n == 0) { o[i] = 0; continue; }
            ld ym = i > 0 ? y[i - 1] : 0, yp = i + 1 < T ? y[i + 1] : 0;
            o[i] = a * (ym - yp) / (2 * (ld)n);
        }
    }
    // I^j [p y]
    std::vector<ld> Ip(int j, const Laurent &p, const std::vector<ld> &y) const
    {
        int deg = -1;
        for (auto &kv : p.c) {
            if (kv.second == 0) continue;
            if (kv.first < 0) throw std::runtime_error("operator coefficient is not polynomial");
            deg = std::max(deg, kv.first);
        }
        std::vector<ld> acc(T, 0), tmp(T);
        if (deg >= 0) {
            auto coef = [&](int e) { auto it = p.c.find(e); return it == p.c.end() ? (ld)0 : it->second; };
            for (int i = 0; i < T; i++) acc[i] = coef(deg) * y[i];
            for (int e = deg - 1; e >= 0; e--) {
                mulr(acc, tmp);
                ld ce = coef(e);
                for (int i = 0; i < T; i++) acc[i] = tmp[i] + ce * y[i];
            }
        }
        for (int q = 0; q < j; q++) { integ(acc, tmp); acc.swap(tmp); }
        return acc;
    }
    // I^j [p D^k y]
    std::vector<ld> term(int j, const Laurent &p, int k, const std::vector<ld> &y) const
    {
        if (k == 0) return Ip(j, p, y);
        if (j < k) throw std::runtime_error("operator is not integrable");
        std::vector<ld> o = term(j - 1, p, k - 1, y), t2 = term(j, p.deriv(), k - 1, y);
        for (int i = 0; i < T; i++) o[i] -= t2[i];
        return o;
    }
};

// unfolded extended rows: ext[n][K + d] = row n, column n + d
struct ExtTab {
    int N = 0, K = 0;
    std::vector<ld> e;
    ld at(int n, int d) const { return (d < -K || d > K) ? 0 : e[(size_t)n * (2 * K + 1) + K + d]; }
};

inline ExtTab make_ext(const DiffOp &op, int J, int N, int K, ld a, ld b)
{
    ExtTab t;
    t.N = N;
    t.K = K;
    t.e.assign((size_t)N * (2 * K + 1), 0);
    SeqOps s;
    s.E = K + 24;
    s.L = N + K + 24;
    s.T = s.E + s.L;
    s.a = a;
    s.b = b;
    std::vector<ld> e(s.T);
    for (int col = -K; col < N + K; col++) {
        std::fill(e.begin(), e.end(), 0);
        e[col + s.E] = 1;
        std::vector<ld> acc(s.T, 0);
        for (int k = 0; k < 8; k++) {
            if (op.a[k].zero()) continue;
            std::vector<ld> o = s.term(J, op.a[k], k, e);
            for (int i = 0; i < s.T; i++) acc[i] += o[i];
        }
        for (int n = 0; n < N; n++) {
            int d = col - n;
            if (d >= -K && d <= K) t.e[(size_t)n * (2 * K + 1) + K + d] = acc[n + s.E];
        }
    }
    return t;
}

// reference stencil (length 2K+1, index nb <-> column idx+K-nb) of
// sum_t coef[t] * tab[t] at row idx, mirrored like mirror_stencil
inline std::vector<ld> stencil(int idx, int K, const std::vector<const ExtTab *> &tabs, const std::vector<ld> &coef)
{
    std::vector<ld> st(2 * K + 1, 0);
    for (int nb = 0; nb < 2 * K + 1; nb++) {
        int d = K - nb;
        ld v = 0;
        for (size_t t = 0; t < tabs.size(); t++) v += coef[t] * tabs[t]->at(idx, d);
        st[nb] = v;
    }
    if (idx < K)
        for (int n = 1; n <= K - idx; n++) st[K - (n - idx)] += st[K - (-n - idx)];
    return st;
}

// The operator library of the integration method (m-independent parts).
struct OpLib {
    ExtTab I2R2, I2R1, L2_0, L2_2, R2, I1R1, I4R4, I4R3, LR_0, LR_2, LR2_0, LR2_2, LR2_4;
    void build(int N, ld r_cmb, ld r_icb)
    {
        ld a = (r_cmb - r_icb) / 2, b = (r_cmb + r_icb) / 2;
        auto poly = [](Laurent p) { DiffOp d; d.a[0] = p; return d; };
        I2R2 = make_ext(poly(Laurent::mono(1, 2)), 2, N, 4, a, b);
        I2R1 = make_ext(poly(Laurent::mono(1, 1)), 2, N, 3, a, b);
        // r^2 Delta_m = r^2 D^2 + r D - m^2
        DiffOp l0, l2;
        l0.a[2] = Laurent::mono(1, 2);
        l0.a[1] = Laurent::mono(1, 1);
        l2.a[0] = Laurent::mono(-1, 0);
        L2_0 = make_ext(l0, 2, N, 2, a, b);
        L2_2 = make_ext(l2, 2, N, 2, a, b);
        R2 = make_ext(poly(Laurent::mono(1, 2)), 0, N, 2, a, b);
        I1R1 = make_ext(poly(Laurent::mono(1, 1)), 1, N, 2, a, b);
        I4R4 = make_ext(poly(Laurent::mono(1, 4)), 4, N, 8, a, b);
        I4R3 = make_ext(poly(Laurent::mono(1, 3)), 4, N, 7, a, b);
        // omega(psi) = -h^2 psi'' - (rc^2/r - 6 r) psi' + (6 + m^2 h^2/r^2) psi
        Laurent h2 = Laurent::mono(r_cmb * r_cmb, 0) + Laurent::mono(-1, 2);
        DiffOp Lr0, Lr2;
        Lr0.a[2] = h2 * (ld)-1;
        Lr0.a[1] = Laurent::mono(-r_cmb * r_cmb, -1) + Laurent::mono(6, 1);
        Lr0.a[0] = Laurent::mono(6, 0);
        Lr2.a[0] = h2 * Laurent::mono(1, -2);
        Laurent r4 = Laurent::mono(1, 4);
        LR_0 = make_ext(Lr0.scaled(r4), 4, N, 8, a, b);
        LR_2 = make_ext(Lr2.scaled(r4), 4, N, 8, a, b);
        // r^4 Delta_m omega = (r^4 D^2 + r^3 D) omega - m^2 r^2 omega
        DiffOp Dm;
        Dm.a[2] = Laurent::mono(1, 4);
        Dm.a[1] = Laurent::mono(1, 3);
        DiffOp A0 = Dm.compose(Lr0), A2 = Dm.compose(Lr2);
        Laurent mr2 = Laurent::mono(-1, 2);
        DiffOp B2 = Lr0.scaled(mr2), B4 = Lr2.scaled(mr2);
        LR2_0 = make_ext(A0, 4, N, 8, a, b);
        LR2_2 = make_ext(A2 + B2, 4, N, 8, a, b);
        LR2_4 = make_ext(B4, 4, N, 8, a, b);
    }
};

} // namespace pzh
