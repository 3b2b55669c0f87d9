// Device code of libpizza_amd: CDNA4 (gfx950) kernels of the IMEX step.
// Included once by pz_api.hip.
#pragma once
#include "pz_internal.h"

// ---------------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ cd mk(double a, double b) { return make_double2(a, b); }
__device__ __forceinline__ cd cadd(cd a, cd b) { return mk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cd csub(cd a, cd b) { return mk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cd cmul(cd a, cd b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ cd cscale(double s, cd a) { return mk(s * a.x, s * a.y); }
__device__ __forceinline__ cd cconj(cd a) { return mk(a.x, -a.y); }
__device__ __forceinline__ cd cdivr(cd a, double s) { return mk(a.x / s, a.y / s); }
// (0 + i s) * a
__device__ __forceinline__ cd cimul(double s, cd a) { return mk(-s * a.y, s * a.x); }
__device__ __forceinline__ cd cdiv(cd a, cd b)
{
    // Smith's algorithm
    if (fabs(b.x) >= fabs(b.y)) {
        double r = b.y / b.x, d = b.x + b.y * r;
        return mk((a.x + a.y * r) / d, (a.y - a.x * r) / d);
    } else {
        double r = b.x / b.y, d = b.y + b.x * r;
        return mk((a.x * r + a.y) / d, (a.y * r - a.x) / d);
    }
}
// matrix element times vector element: real or complex matrix
__device__ __forceinline__ cd mmul(double a, cd v) { return mk(a * v.x, a * v.y); }
__device__ __forceinline__ cd mmul(cd a, cd v) { return cmul(a, v); }
__device__ __forceinline__ cd mdiv(cd v, double a) { return mk(v.x / a, v.y / a); }
__device__ __forceinline__ cd mdiv(cd v, cd a) { return cdiv(v, a); }
__device__ __forceinline__ double mabs1(double a) { return fabs(a); }
__device__ __forceinline__ double mabs1(cd a) { return fabs(a.x) + fabs(a.y); }
__device__ __forceinline__ bool mnz(double a) { return a != 0.0; }
__device__ __forceinline__ bool mnz(cd a) { return a.x != 0.0 || a.y != 0.0; }
__device__ __forceinline__ double mrecip(double a) { return 1.0 / a; }
__device__ __forceinline__ cd mrecip(cd a) { return cdiv(mk(1.0, 0.0), a); }
__device__ __forceinline__ double mscale(double a, double s) { return a * s; }
__device__ __forceinline__ cd mscale(cd a, double s) { return mk(a.x * s, a.y * s); }
__device__ __forceinline__ double mmm(double a, double b) { return a * b; }
__device__ __forceinline__ cd mmm(cd a, cd b) { return cmul(a, b); }
__device__ __forceinline__ double mneg(double a) { return -a; }
__device__ __forceinline__ cd mneg(cd a) { return mk(-a.x, -a.y); }
__device__ __forceinline__ double madd(double a, double b) { return a + b; }
template <class M> __device__ __forceinline__ M mzero();
template <> __device__ __forceinline__ double mzero<double>() { return 0.0; }
template <> __device__ __forceinline__ cd mzero<cd>() { return make_double2(0.0, 0.0); }
__device__ __forceinline__ cd madd(cd a, cd b) { return cadd(a, b); }

// wave-uniform lane read of a double2 (lane index must be uniform)
__device__ __forceinline__ double rl(double v, int lane)
{
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ cd rlc(cd v, int lane) { return mk(rl(v.x, lane), rl(v.y, lane)); }

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Element access through a generic pointer known to address LDS, retyped so
// that it compiles to ds_read/ds_write instead of flat loads and stores.
typedef __attribute__((address_space(3))) double lds_double;
__device__ __forceinline__ cd lds_load(const cd *q, int i)
{
    const lds_double *d = (const lds_double *)q;
    return mk(d[2 * i], d[2 * i + 1]);
}
__device__ __forceinline__ void lds_store(cd *q, int i, cd v)
{
    lds_double *d = (lds_double *)q;
    d[2 * i] = v.x;
    d[2 * i + 1] = v.y;
}

// Loads through a generic pointer known to address global memory, retyped
// so that they compile to global_load (which counts only against vmcnt)
// instead of flat loads, whose waits also stall on pending LDS accesses.
typedef double pz_d2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const pz_d2 g_d2;
typedef __attribute__((address_space(1))) const double g_double;
typedef __attribute__((address_space(1))) const int g_int;
__device__ __forceinline__ double gld(const double *q, size_t i) { return ((g_double *)q)[i]; }
__device__ __forceinline__ cd gld(const cd *q, size_t i)
{
    const pz_d2 v = ((g_d2 *)q)[i];
    return mk(v.x, v.y);
}
// two consecutive doubles at q (16-byte aligned)
__device__ __forceinline__ cd gld2(const double *q)
{
    const pz_d2 v = *(g_d2 *)q;
    return mk(v.x, v.y);
}

// Up to four vector pointers passed by value; indexing by selects keeps
// them in registers (a private array indexed at run time would live in
// scratch memory and turn LDS accesses into flat ones).
template <class T>
struct VecSet {
    T *v0 = nullptr, *v1 = nullptr, *v2 = nullptr, *v3 = nullptr;
    int n = 0;
    __device__ VecSet() {}
    // unused slots repeat an earlier pointer, so every slot stays in the
    // address space of the vectors (no null among the selected pointers)
    __device__ explicit VecSet(T *a, T *b = nullptr, T *c = nullptr, T *d = nullptr)
        : v0(a), v1(b ? b : a), v2(c ? c : a), v3(d ? d : a), n(d ? 4 : c ? 3 : b ? 2 : a ? 1 : 0) {}
    __device__ __forceinline__ void push(T *q)
    {
        v0 = n == 0 ? q : v0;
        v1 = n <= 1 ? q : v1;
        v2 = n <= 2 ? q : v2;
        v3 = n <= 3 ? q : v3;
        n++;
    }
    __device__ __forceinline__ T *operator[](int f) const { return f == 0 ? v0 : f == 1 ? v1 : f == 2 ? v2 : v3; }
};

// ---------------------------------------------------------------------------
// Stockham mixed-radix FFT in LDS for ntr transforms of length n
// (dir = -1 forward, +1 backward, unnormalised).  Result pointer returned.
// ---------------------------------------------------------------------------
__device__ __forceinline__ cd twd(const PzDev &p, int k, int dir)
{
    cd w = p.twid[k];
    return dir < 0 ? w : mk(w.x, -w.y);
}

// Radix-R butterfly of a Stockham stage (R = 2, 3, 4, 8): the inputs
// x[j + r nb] times the stage twiddles tw(r tstep), outputs v[r] (to be
// stored R outputs apart by Ns)
template <int R, class TW>
__device__ __forceinline__ void stockham_bfly(const cd *x, int j, int nb, int tstep, int dir, TW tw, cd (&v)[R])
{
    if constexpr (R == 2) {
        const cd v0 = x[j], v1 = cmul(x[j + nb], tw(tstep));
        v[0] = cadd(v0, v1);
        v[1] = csub(v0, v1);
    } else if constexpr (R == 4) {
        const cd v0 = x[j];
        const cd v1 = cmul(x[j + nb], tw(tstep));
        const cd v2 = cmul(x[j + 2 * nb], tw(2 * tstep));
        const cd v3 = cmul(x[j + 3 * nb], tw(3 * tstep));
        const cd a0 = cadd(v0, v2), b0 = csub(v0, v2), c0 = cadd(v1, v3);
        const cd d0 = cimul((double)dir, csub(v1, v3));
        v[0] = cadd(a0, c0);
        v[1] = cadd(b0, d0);
        v[2] = csub(a0, c0);
        v[3] = csub(b0, d0);
    } else if constexpr (R == 8) {
        // radix 8 as 2 x 4: a_r = v_r + v_{r+4} gives the even outputs,
        // (v_r - v_{r+4}) w8^r the odd ones, w8 = exp(dir i pi/4)
        const cd v0 = x[j], v1 = cmul(x[j + nb], tw(tstep)), v2 = cmul(x[j + 2 * nb], tw(2 * tstep));
        const cd v3 = cmul(x[j + 3 * nb], tw(3 * tstep)), v4 = cmul(x[j + 4 * nb], tw(4 * tstep));
        const cd v5 = cmul(x[j + 5 * nb], tw(5 * tstep)), v6 = cmul(x[j + 6 * nb], tw(6 * tstep));
        const cd v7 = cmul(x[j + 7 * nb], tw(7 * tstep));
        const double h = 0.70710678118654752440, sd = (double)dir;
        const cd a0 = cadd(v0, v4), a1 = cadd(v1, v5), a2 = cadd(v2, v6), a3 = cadd(v3, v7);
        const cd b0 = csub(v0, v4), c1 = csub(v1, v5), c2 = csub(v2, v6), c3 = csub(v3, v7);
        const cd b1 = mk(h * (c1.x - sd * c1.y), h * (c1.y + sd * c1.x));
        const cd b2 = cimul(sd, c2);
        const cd b3 = mk(h * (-c3.x - sd * c3.y), h * (-c3.y + sd * c3.x));
        {
            const cd s02 = cadd(a0, a2), d02 = csub(a0, a2), s13 = cadd(a1, a3), d13 = cimul(sd, csub(a1, a3));
            v[0] = cadd(s02, s13);
            v[2] = cadd(d02, d13);
            v[4] = csub(s02, s13);
            v[6] = csub(d02, d13);
        }
        {
            const cd s02 = cadd(b0, b2), d02 = csub(b0, b2), s13 = cadd(b1, b3), d13 = cimul(sd, csub(b1, b3));
            v[1] = cadd(s02, s13);
            v[3] = cadd(d02, d13);
            v[5] = csub(s02, s13);
            v[7] = csub(d02, d13);
        }
    } else if constexpr (R == 6) {
        // radix 6 as 2 x 3: the DFT3s E of the even and O of the odd taps,
        // X_k = E_k + w6^k O_k, X_{k+3} = E_k - w6^k O_k, w6 = exp(dir i pi/3)
        const double s3 = 0.86602540378443864676, sd = (double)dir;
        cd e[3], o[3];
        const cd v0 = x[j], v1 = cmul(x[j + nb], tw(tstep)), v2 = cmul(x[j + 2 * nb], tw(2 * tstep));
        const cd v3 = cmul(x[j + 3 * nb], tw(3 * tstep)), v4 = cmul(x[j + 4 * nb], tw(4 * tstep));
        const cd v5 = cmul(x[j + 5 * nb], tw(5 * tstep));
        auto dft3 = [&](cd a0, cd a1, cd a2, cd(&r)[3]) {
            const cd t1 = cadd(a1, a2), t2 = cimul(sd * s3, csub(a1, a2));
            const cd m0 = mk(a0.x - 0.5 * t1.x, a0.y - 0.5 * t1.y);
            r[0] = cadd(a0, t1);
            r[1] = cadd(m0, t2);
            r[2] = csub(m0, t2);
        };
        dft3(v0, v2, v4, e);
        dft3(v1, v3, v5, o);
        const cd o1 = cmul(mk(0.5, sd * s3), o[1]), o2 = cmul(mk(-0.5, sd * s3), o[2]);
        v[0] = cadd(e[0], o[0]);
        v[3] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[4] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[5] = csub(e[2], o2);
    } else {
        static_assert(R == 3, "radix");
        const double s3 = 0.86602540378443864676;
        const cd v0 = x[j];
        const cd v1 = cmul(x[j + nb], tw(tstep));
        const cd v2 = cmul(x[j + 2 * nb], tw(2 * tstep));
        const cd t1 = cadd(v1, v2);
        const cd t2 = cimul((double)dir * s3, csub(v1, v2));
        const cd m0 = mk(v0.x - 0.5 * t1.x, v0.y - 0.5 * t1.y);
        v[0] = cadd(v0, t1);
        v[1] = cadd(m0, t2);
        v[2] = csub(m0, t2);
    }
}

template <class TW>
__device__ cd *fft_stockham(cd *a, cd *b, int n, int ntr, int dir, const int *radices, int nrad, TW tw)
{
    int Ns = 1;
    for (int s = 0; s < nrad; s++) {
        const int R = radices[s];
        const int nb = n / R;
        if (R != 2 && R != 3 && R != 4 && R != 8) {
            // other radices: a direct DFT of length R per output, one output
            // per item so that all n outputs of a stage run in parallel.  The
            // stage twiddle and the DFT kernel of tap r combine into
            // exp(dir 2 pi i (r tstep + r k n/R) / n) = z^r with
            // z = exp(dir 2 pi i (tstep + k n/R) / n), tstep + k n/R < n, so
            // the output is the polynomial sum_r x_r z^r, evaluated by
            // Horner's rule: one table entry and R complex FMAs per output.
            const int q = n / R;
            for (int t = threadIdx.x; t < n * ntr; t += blockDim.x) {
                const int tr = t / n, rem = t - tr * n, k = rem / nb, j = rem - k * nb;
                const cd *x = a + tr * n;
                const int jm = j % Ns, tstep = (n / (Ns * R)) * jm;
                const cd z = tw(tstep + k * q);
                cd acc = x[j + (R - 1) * nb];
                for (int r = R - 2; r >= 0; r--) {
                    const cd v = x[j + r * nb];
                    acc = mk(fma(acc.x, z.x, fma(-acc.y, z.y, v.x)), fma(acc.x, z.y, fma(acc.y, z.x, v.y)));
                }
                b[tr * n + (j / Ns) * Ns * R + jm + k * Ns] = acc;
            }
            __syncthreads();
            cd *t = a;
            a = b;
            b = t;
            Ns *= R;
            continue;
        }
        const int tot = nb * ntr;
        for (int t = threadIdx.x; t < tot; t += blockDim.x) {
            const int tr = t / nb, j = t - tr * nb, jm = j % Ns;
            const int tstep = (n / (Ns * R)) * jm, idxD = (j / Ns) * Ns * R + jm;
            const cd *x = a + tr * n;
            cd *y = b + tr * n;
            if (R == 2) {
                cd v[2];
                stockham_bfly<2>(x, j, nb, tstep, dir, tw, v);
                y[idxD] = v[0]; y[idxD + Ns] = v[1];
            } else if (R == 4) {
                cd v[4];
                stockham_bfly<4>(x, j, nb, tstep, dir, tw, v);
#pragma unroll
                for (int r = 0; r < 4; r++) y[idxD + r * Ns] = v[r];
            } else if (R == 8) {
                cd v[8];
                stockham_bfly<8>(x, j, nb, tstep, dir, tw, v);
#pragma unroll
                for (int r = 0; r < 8; r++) y[idxD + r * Ns] = v[r];
            } else if (R == 3) {
                cd v[3];
                stockham_bfly<3>(x, j, nb, tstep, dir, tw, v);
                y[idxD] = v[0]; y[idxD + Ns] = v[1]; y[idxD + 2 * Ns] = v[2];
            }
        }
        __syncthreads();
        cd *t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    return a;
}

// The same transforms in place, for the long azimuthal grids: every stage
// computes its outputs into registers, then a barrier, the stores, and a
// barrier, so one LDS column per transform is enough.  A thread takes at
// most IT items of a stage (fft_inplace_fits checks the plan on the host):
// a butterfly of radix 2, 3, 4 or 6, or four consecutive outputs of another
// radix (8 included), which share their input reads (Horner's rule as
// above).
template <int R, int IT, class TW>
__device__ __forceinline__ void inplace_bfly(cd *a, int n, int ntr, int Ns, int dir, TW tw)
{
    const int nb = n / R, tot = ntr * nb;
    cd y[IT][R];
    int d[IT];
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int t = threadIdx.x + it * blockDim.x;
        d[it] = -1;
        if (t < tot) {
            const int tr = t / nb, j = t - tr * nb, jm = j % Ns;
            stockham_bfly<R>(a + tr * n, j, nb, (n / (Ns * R)) * jm, dir, tw, y[it]);
            d[it] = tr * n + (j / Ns) * Ns * R + jm;
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; it++)
        if (d[it] >= 0)
#pragma unroll
            for (int r = 0; r < R; r++) a[d[it] + r * Ns] = y[it][r];
    __syncthreads();
}
// one Horner step a z + v
__device__ __forceinline__ cd hstep(cd a, cd z, cd v)
{
    return mk(fma(a.x, z.x, fma(-a.y, z.y, v.x)), fma(a.x, z.y, fma(a.y, z.x, v.y)));
}
template <int IT, int G, class TW>
__device__ __forceinline__ void inplace_prime(cd *a, int n, int ntr, int Ns, int R, TW tw)
{
    const int nb = n / R, ng = (R + G - 1) / G, tot = ntr * nb * ng;
    cd y[IT][G];
    int d[IT], cnt[IT];
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int t = threadIdx.x + it * blockDim.x;
        cnt[it] = 0;
        d[it] = 0;
        if (t < tot) {
            const int u = t / ng, g = t - u * ng, tr = u / nb, j = u - tr * nb;
            const int jm = j % Ns, tstep = (n / (Ns * R)) * jm, k0 = G * g;
            const cd *x = a + tr * n + j;
            cd z[G], acc[G];
            const cd top = x[(R - 1) * nb];
#pragma unroll
            for (int c = 0; c < G; c++) {
                z[c] = tw(tstep + min(k0 + c, R - 1) * nb);
                acc[c] = top;
            }
            for (int r = R - 2; r >= 0; r--) {
                const cd v = x[r * nb];
#pragma unroll
                for (int c = 0; c < G; c++) acc[c] = hstep(acc[c], z[c], v);
            }
#pragma unroll
            for (int c = 0; c < G; c++) y[it][c] = acc[c];
            d[it] = tr * n + (j / Ns) * Ns * R + jm + k0 * Ns;
            cnt[it] = min(G, R - k0);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; it++)
#pragma unroll
        for (int c = 0; c < G; c++)
            if (c < cnt[it]) a[d[it] + c * Ns] = y[it][c];
    __syncthreads();
}
// The first stage of a plan (no twiddles) with an odd prime R, 5 <= R <=
// 31: the R-point DFT by conjugate pairs.  With s_j = x_j + x_{R-j}, d_j =
// x_j - x_{R-j} and w^{jk} = c + i sigma (the direction's sign in sigma),
// X_k = x_0 + A_k + i B_k and X_{R-k} = x_0 + A_k - i B_k, where A_k = sum_j
// c_{jk} s_j and B_k = sum_j sigma_{jk} d_j (j, k = 1 .. h, h = (R-1)/2):
// two real-by-complex products per term and pair instead of two complex
// Horner steps per output.  om[(j-1) PZ_SYM_ROW + k-1] = w^{jk mod R} (a
// row per j, filled here); an item takes P consecutive pairs of one
// butterfly (the first item also X_0 = x_0 + sum s_j).
#define PZ_SYM_PAIRS 5
#define PZ_SYM_ROW 16
#define PZ_FFT_OM_MAX (15 * PZ_SYM_ROW)
__host__ __device__ inline bool sym_first_stage(int R) { return R >= 5 && (R & 1) && R <= 31; }
template <int P, class TW>
__device__ __forceinline__ void inplace_prime_first(cd *a, int n, int ntr, int R, cd *om, TW tw)
{
    const int nb = n / R, h = (R - 1) >> 1, ng = (h + P - 1) / P, tot = ntr * nb * ng;
    for (int e = threadIdx.x; e < h * PZ_SYM_ROW; e += blockDim.x) {
        const int j = e / PZ_SYM_ROW + 1, k = e - (j - 1) * PZ_SYM_ROW + 1;
        om[e] = tw(((j * k) % R) * nb);
    }
    __syncthreads();
    cd y[2 * P + 1];
    int base = 0, kl = 1, cnt = 0;
    const int t = threadIdx.x;
    if (t < tot) {
        const int u = t / ng, g = t - u * ng, tr = u / nb, j = u - tr * nb;
        const cd *xp = a + tr * n + j, *xm = xp + R * nb;
        const cd x0 = *xp;
        kl = 1 + P * g;
        const cd *w = om + (kl - 1);
        cd A[P], B[P];
#pragma unroll
        for (int c = 0; c < P; c++) { A[c] = mk(0.0, 0.0); B[c] = mk(0.0, 0.0); }
        cd sum = x0;
        for (int jj = 1; jj <= h; jj++) {
            xp += nb;
            xm -= nb;
            const cd vp = *xp, vm = *xm;
            const cd sj = cadd(vp, vm), dj = csub(vp, vm);
            sum = cadd(sum, sj);
#pragma unroll
            for (int c = 0; c < P; c++) {
                const cd wc = w[c];
                A[c] = mk(fma(wc.x, sj.x, A[c].x), fma(wc.x, sj.y, A[c].y));
                B[c] = mk(fma(wc.y, dj.x, B[c].x), fma(wc.y, dj.y, B[c].y));
            }
            w += PZ_SYM_ROW;
        }
#pragma unroll
        for (int c = 0; c < P; c++) {
            y[2 * c] = mk(x0.x + A[c].x - B[c].y, x0.y + A[c].y + B[c].x);
            y[2 * c + 1] = mk(x0.x + A[c].x + B[c].y, x0.y + A[c].y - B[c].x);
        }
        y[2 * P] = g == 0 ? sum : mk(0.0, 0.0);
        base = tr * n + j * R;
        cnt = min(P, h + 1 - kl);
        if (g == 0) cnt = -1 - cnt; // X_0 as well
    }
    __syncthreads();
    const int np = cnt < 0 ? -1 - cnt : cnt;
#pragma unroll
    for (int c = 0; c < P; c++)
        if (c < np) {
            a[base + kl + c] = y[2 * c];
            a[base + R - kl - c] = y[2 * c + 1];
        }
    if (cnt < 0) a[base] = y[2 * P];
    __syncthreads();
}
// om: LDS for PZ_FFT_OM_MAX entries, used when the plan starts with an odd
// prime 5 <= R <= 31 (one item per thread there: fft_inplace_fits)
template <class TW>
__device__ __forceinline__ void fft_inplace(cd *a, int n, int ntr, int dir, const int *radices, int nrad, TW tw, cd *om)
{
    int Ns = 1;
    for (int s = 0; s < nrad; s++) {
        const int R = radices[s];
        if (s == 0 && sym_first_stage(R)) {
            inplace_prime_first<PZ_SYM_PAIRS>(a, n, ntr, R, om, tw);
            Ns *= R;
            continue;
        }
        switch (R) {
        case 2: inplace_bfly<2, 2>(a, n, ntr, Ns, dir, tw); break;
        case 3: inplace_bfly<3, 2>(a, n, ntr, Ns, dir, tw); break;
        case 4: inplace_bfly<4, 1>(a, n, ntr, Ns, dir, tw); break;
        case 6: inplace_bfly<6, 2>(a, n, ntr, Ns, dir, tw); break;
        default: inplace_prime<3, 4>(a, n, ntr, Ns, R, tw); break; // up to three items of four outputs
        }
        Ns *= R;
    }
}
// host check: every stage of the plan gives each of nthreads threads at
// most the items that fft_inplace holds in registers
inline bool fft_inplace_fits(int n, int ntr, const int *radices, int nrad, int nthreads)
{
    for (int s = 0; s < nrad; s++) {
        const int R = radices[s], nb = n / R;
        long items = (long)ntr * nb, cap = R == 4 ? 1 : 2; // radix 2, 3, 6: two butterflies
        if (s == 0 && sym_first_stage(R)) {
            items *= ((R - 1) / 2 + PZ_SYM_PAIRS - 1) / PZ_SYM_PAIRS;
            cap = 1;
        } else if (R != 2 && R != 3 && R != 4 && R != 6) {
            items *= (R + 3) / 4; // four outputs an item, three items a thread
            cap = 3;
        }
        if (items > cap * nthreads) return false;
    }
    return true;
}
// the plan of the in-place transforms: the largest odd prime up to 31
// first (its stage has no twiddles), a radix 2 and a radix 3 merged into
// one radix-6 stage, the other radices in their order; returns the number
// of stages
inline int inplace_plan(int *radices, int nrad)
{
    int i2 = -1, i3 = -1;
    for (int s = 0; s < nrad; s++) {
        if (radices[s] == 2 && i2 < 0) i2 = s;
        if (radices[s] == 3 && i3 < 0) i3 = s;
    }
    if (i2 >= 0 && i3 >= 0) {
        radices[std::min(i2, i3)] = 6;
        for (int s = std::max(i2, i3); s + 1 < nrad; s++) radices[s] = radices[s + 1];
        nrad--;
    }
    int best = -1;
    for (int s = 0; s < nrad; s++)
        if (sym_first_stage(radices[s]) && (best < 0 || radices[s] > radices[best])) best = s;
    if (best > 0) {
        const int R = radices[best];
        for (int s = best; s > 0; s--) radices[s] = radices[s - 1];
        radices[0] = R;
    }
    return nrad;
}

// azimuthal FFT of length n_phi (twiddles exp(-2 pi i k / n_phi) in p.twid)
__device__ cd *fft_lds(const PzDev &p, cd *a, cd *b, int n, int ntr, int dir)
{
    return fft_stockham(a, b, n, ntr, dir, p.radices, p.nrad, [&](int k) { return twd(p, k, dir); });
}

// FFT of length L = 2 n_r - 2 with twiddles exp(dir 2 pi i k / L) =
// (cosT[k], dir sinT[k]) read from the Chebyshev tables
__device__ cd *fft_cheb(const PzDev &p, cd *a, cd *b, int ntr, int dir, const double *cosT, const double *sinT)
{
    const double sd = (double)dir;
    return fft_stockham(a, b, p.L, ntr, dir, p.radicesL, p.nradL,
                        [&](int k) { return mk(cosT[k], sd * sinT[k]); });
}

// ---------------------------------------------------------------------------
// FFTs of length L = 64 P held in registers, one wave per transform
// (waves loop over the transforms): lane l holds x[j1 + P l],
// j1 < P.  With j = j1 + P j2 and k = k2 + 64 k1,
//   X[k] = sum_j1 w_P^(j1 k1) w_L^(j1 k2) sum_j2 w_64^(j2 k2) x[j1 + P j2]:
// the inner 64-point DFTs run across the lanes as six radix-2
// decimation-in-frequency stages with xor-shuffles (lane l then holds
// k2 = bitreverse(l)), the outer P-point DFT in registers.  No barrier:
// ld(tr, j) supplies the input, st(tr, k, X) takes each output once;
// tw(q) = exp(-2 pi i q / L) for the forward transform, its conjugate for
// the backward one.  All twiddles of a lane are read before the first
// stage, so their latencies overlap.
// x of lane l ^ H within the wave, for a compile-time H: DPP quad
// permutations for H = 1, 2, swizzles for H = 4, 8, 16, a permute for 32
template <int H>
__device__ __forceinline__ int xor_lane(int x)
{
    if constexpr (H == 1) return __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false); // quad_perm(1,0,3,2)
    else if constexpr (H == 2) return __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false); // quad_perm(2,3,0,1)
    else if constexpr (H <= 16) return __builtin_amdgcn_ds_swizzle(x, 0x1F | (H << 10));  // bit mode, xor H
    else return __builtin_amdgcn_ds_bpermute(((int)(threadIdx.x & 63) ^ H) << 2, x);
}
template <int H>
__device__ __forceinline__ double xor_lane_d(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = xor_lane<H>((int)b), hi = xor_lane<H>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// one radix-2 decimation-in-frequency stage of span H across the lanes
template <int P, int H>
__device__ __forceinline__ void dif_stage(cd (&v)[P], cd w, int lane)
{
    const bool up = (lane & H) != 0;
#pragma unroll
    for (int j1 = 0; j1 < P; j1++) {
        const cd o = mk(xor_lane_d<H>(v[j1].x), xor_lane_d<H>(v[j1].y));
        v[j1] = up ? cmul(csub(o, v[j1]), w) : cadd(v[j1], o);
    }
}

template <int P, class LD, class ST, class TW>
__device__ __forceinline__ void fft_wave(int ntr, LD ld, ST st, TW tw)
{
    const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    const int k2 = (int)(__brev((unsigned)lane) >> 26);
    cd ws[6], wk[P > 1 ? P - 1 : 1], wp[P > 1 ? P : 1];
#pragma unroll
    for (int s = 0; s < 6; s++) ws[s] = tw((lane & ((1 << s) - 1)) * (32 >> s) * P); // W_2h^t, h = 2^s
#pragma unroll
    for (int j1 = 1; j1 < P; j1++) wk[j1 - 1] = tw(j1 * k2);
#pragma unroll
    for (int q = 0; q < P; q++) wp[P > 1 ? q : 0] = tw(q * 64); // w_P^q
    for (int tr = threadIdx.x >> 6; tr < ntr; tr += nw) {
        cd v[P];
#pragma unroll
        for (int j1 = 0; j1 < P; j1++) v[j1] = ld(tr, j1 + P * lane);
        dif_stage<P, 32>(v, ws[5], lane);
        dif_stage<P, 16>(v, ws[4], lane);
        dif_stage<P, 8>(v, ws[3], lane);
        dif_stage<P, 4>(v, ws[2], lane);
        dif_stage<P, 2>(v, ws[1], lane);
        dif_stage<P, 1>(v, ws[0], lane);
#pragma unroll
        for (int j1 = 1; j1 < P; j1++) v[j1] = cmul(v[j1], wk[j1 - 1]);
#pragma unroll
        for (int k1 = 0; k1 < P; k1++) {
            cd acc = v[0];
#pragma unroll
            for (int j1 = 1; j1 < P; j1++) acc = cadd(acc, cmul(v[j1], wp[(j1 * k1) % P]));
            st(tr, k2 + 64 * k1, acc);
        }
    }
}

template <class LD, class ST, class TW>
__device__ __forceinline__ void fft_wave_p(int P, int ntr, LD ld, ST st, TW tw)
{
    switch (P) {
    case 1: fft_wave<1>(ntr, ld, st, tw); break;
    case 2: fft_wave<2>(ntr, ld, st, tw); break;
    case 3: fft_wave<3>(ntr, ld, st, tw); break;
    default: fft_wave<4>(ntr, ld, st, tw); break;
    }
}

// The azimuthal transforms of an n_phi = 64 P grid, P in {2, 3, 4, 6}
template <class LD, class ST, class TW>
__device__ __forceinline__ void fft_wave_phi(int P, int ntr, LD ld, ST st, TW tw)
{
    switch (P) {
    case 2: fft_wave<2>(ntr, ld, st, tw); break;
    case 3: fft_wave<3>(ntr, ld, st, tw); break;
    case 4: fft_wave<4>(ntr, ld, st, tw); break;
    default: fft_wave<6>(ntr, ld, st, tw); break;
    }
}
__device__ __forceinline__ int wave_fft_p_phi(int n)
{
    const int P = n >> 6;
    return (n & 63) == 0 && (P == 2 || P == 3 || P == 4 || P == 6) ? P : 0;
}

// P of the register FFTs for this run's Chebyshev length, 0 when L is not
// 64, 128, 192 or 256 (n_r 33, 65, 97, 129)
__device__ __forceinline__ int wave_fft_p(const PzDev &p)
{
    return (p.L & 63) == 0 && p.L <= 256 ? p.L >> 6 : 0;
}

__device__ __forceinline__ double cc22real(cd a, cd b, int m)
{
    return m == 0 ? a.x * b.x : 2.0 * (a.x * b.x + a.y * b.y);
}

__device__ double block_reduce(double v, double *scratch, int op /*0 sum, 1 max*/)
{
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    for (int o = 32; o > 0; o >>= 1) {
        double u = __shfl_xor(v, o, 64);
        v = op == 0 ? v + u : fmax(v, u);
    }
    __syncthreads();
    if (lane == 0) scratch[w] = v;
    __syncthreads();
    double r = scratch[0];
    for (int i = 1; i < nw; i++) r = op == 0 ? r + scratch[i] : fmax(r, scratch[i]);
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// Radial loop: rloop::radial_loop (radial_loop.f90:264-392) for one radius
// per workgroup.  Real fields travel in pairs through one complex FFT.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void radial_loop_body(const PzDev &p, int nr, int is, int courant)
{
    extern __shared__ double smem[];
    const int n = p.nphi, nm = p.nm, N = p.N;
    const size_t F = (size_t)nm * N;
    cd *spec = (cd *)smem;           // 5 x nm
    cd *bufa = spec + 5 * nm;        // 3 x n
    cd *bufb = bufa + 3 * n;         // 3 x n
    cd *twl = bufb + 3 * n;          // n twiddles exp(-2 pi i k / n)
    double *red = (double *)(twl + n);
    const int nin = (p.l_heat || p.l_chem) ? (p.l_chem ? 3 : 2) : 2;

    for (int k = threadIdx.x; k < n; k += blockDim.x) twl[k] = p.twid[k];
    for (int i = threadIdx.x; i < nm; i += blockDim.x) {
        size_t q = (size_t)i * N + nr;
        spec[i] = p.us[q];
        spec[nm + i] = p.up[q];
        spec[2 * nm + i] = p.om[q];
        spec[3 * nm + i] = p.l_heat ? p.temp[q] : mk(0, 0);
        spec[4 * nm + i] = p.l_chem ? p.xi[q] : mk(0, 0);
    }
    __syncthreads();
    // Reynolds-stress m-sum (radial_loop.f90:267-278)
    double part = 0.0;
    for (int i = threadIdx.x; i < nm; i += blockDim.x) {
        int m = p.idx2m[i];
        cd usf = m == 0 ? cscale(0.5 * p.ek * p.CorFac * p.ekpump[nr], spec[nm + i]) : spec[i];
        part += cc22real(usf, spec[2 * nm + i], m);
    }
    double usom = block_reduce(part, red, 0);
    // pairs (us,up), (om,T), (xi,0) as Hermitian spectra of z = a + i b
    auto packed = [=](int tr, int k) {
        // a + i b from the m < n_m half, its conjugate from the other
        const bool lo = k < nm, hi = !lo && n - k < nm;
        const int kk = lo ? k : n - k;
        if (!lo && !hi) return mk(0.0, 0.0);
        const cd a = spec[(2 * tr) * nm + kk];
        const cd b = (2 * tr + 1) < 5 ? spec[(2 * tr + 1) * nm + kk] : mk(0.0, 0.0);
        if (k == 0) return mk(a.x, b.x);
        return lo ? mk(a.x - b.y, a.y + b.x) : mk(a.x + b.y, b.x - a.y);
    };
    // register FFTs (one wave per transform, packing as they load) where
    // n_phi = 64 P, else LDS Stockham FFTs of the packed columns
    const int Pw = wave_fft_p_phi(n);
    cd *g;
    if (Pw) {
        fft_wave_phi(Pw, nin, packed, [=](int tr, int k, cd v) { bufb[tr * n + k] = v; },
                     [=](int k) { const cd w = twl[k]; return mk(w.x, -w.y); });
        __syncthreads();
        g = bufb;
    } else {
        for (int t = threadIdx.x; t < nin * n; t += blockDim.x) {
            const int tr = t / n;
            bufa[t] = packed(tr, t - tr * n);
        }
        __syncthreads();
        g = fft_stockham(bufa, bufb, n, nin, +1, p.radices, p.nrad,
                         [&](int k) { const cd w = twl[k]; return mk(w.x, -w.y); });
    }
    cd *o = (g == bufa) ? bufb : bufa;
    // Courant (courant.f90:41-76)
    if (courant) {
        double vr = 0.0, vh = 0.0;
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            cd z = g[j];
            vr = fmax(vr, z.x * z.x);
            vh = fmax(vh, z.y * z.y);
        }
        vr = block_reduce(vr, red, 1);
        vh = block_reduce(vh, red, 1);
        if (threadIdx.x == 0) {
            double cf = p.courfac;
            vr = cf * cf * vr;
            vh = cf * cf * vh;
            double dtrk = 1.0e10, dthk = 1.0e10;
            if (vr != 0.0) dtrk = fmin(dtrk, sqrt(p.delxr2[nr] / vr));
            if (vh != 0.0) dthk = fmin(dthk, sqrt(p.delxh2[nr] / vh));
            p.dtr[nr] = dtrk;
            p.dth[nr] = dthk;
        }
    }
    // nonlinear products (radial_loop.f90:316-325): (upX + i r us X)
    const double rr = p.r[nr];
    const int nout = 1 + (p.l_heat ? 1 : 0) + (p.l_chem ? 1 : 0); // omega, then T and xi as present
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        cd z0 = g[j], z1 = g[n + j];
        double us = z0.x, up = z0.y, om = z1.x, T = z1.y;
        o[j] = mk(up * om, rr * us * om);
        if (p.l_heat) o[n + j] = mk(up * T, rr * us * T);
        if (p.l_chem) {
            double X = g[2 * n + j].x;
            o[(p.l_heat ? 2 : 1) * n + j] = mk(up * X, rr * us * X);
        }
    }
    __syncthreads();
    cd *other = (o == bufa) ? bufb : bufa;
    cd *Z;
    if (Pw) {
        fft_wave_phi(Pw, nout, [=](int tr, int i) { return o[tr * n + i]; },
                     [=](int tr, int k, cd v) { other[tr * n + k] = v; }, [=](int k) { return twl[k]; });
        __syncthreads();
        Z = other;
    } else {
        Z = fft_stockham(o, other, n, nout, -1, p.radices, p.nrad, [&](int k) { return twl[k]; });
    }
    // unpack, normalise and post-scale (radial_loop.f90:350-386)
    const double or1 = p.or1[nr];
    const double fbeta = (1.0 - p.tadvz) * p.beta[nr] * or1;
    for (int t = threadIdx.x; t < nout * nm; t += blockDim.x) {
        int tr = t / nm, i = t - tr * nm;
        cd zk = Z[tr * n + i], zc = cconj(Z[tr * n + (i == 0 ? 0 : n - i)]);
        cd A = mk(0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y));  // spectrum of up*X
        cd B = mk(0.5 * (zk.y - zc.y), -0.5 * (zk.x - zc.x)); // spectrum of r us X
        A = cdivr(A, (double)n);
        B = cdivr(B, (double)n);
        int m = p.idx2m[i];
        size_t q = (size_t)i * N + nr;
        cd cim = mk(0.0, -or1 * (double)m); // -or1*ci*m
        if (tr == 0) {
            p.P_expl[is * F + q] = (i == 0) ? mk(-usom, 0.0) : cmul(cim, A);
            p.dVsOm[q] = B;
        } else if (tr == 1 && p.l_heat) {
            cd v = cmul(cim, A);
            p.T_expl[is * F + q] = mk(v.x - fbeta * B.x, v.y - fbeta * B.y);
            p.dVsT[q] = B;
        } else {
            p.X_expl[is * F + q] = cmul(cim, A);
            p.dVsXi[q] = B;
        }
    }
}

// The radial loop for long azimuthal grids (n_phi too large for the staged
// variant above): the spectra are packed straight from global memory and
// the real fields travel through the complex FFTs one pair at a time, so
// LDS holds four columns of n_phi values (three fields and one scratch).
__device__ __forceinline__ void radial_loop_big_body(const PzDev &p, int nr, int is, int courant)
{
    extern __shared__ double smem[];
    const int n = p.nphi, nm = p.nm, N = p.N;
    const size_t F = (size_t)nm * N;
    cd *s0 = (cd *)smem, *s1 = s0 + n, *s2 = s1 + n, *sc = s2 + n;
    double *red = (double *)(sc + n);
    // without composition the third column holds the twiddle table
    const bool twl_lds = !p.l_chem;
    cd *twl = s2;
    if (twl_lds)
        for (int k = threadIdx.x; k < n; k += blockDim.x) twl[k] = p.twid[k];
    auto fwd = [&](int k) { return twl_lds ? twl[k] : p.twid[k]; };
    auto bwd = [&](int k) {
        const cd w = twl_lds ? twl[k] : p.twid[k];
        return mk(w.x, -w.y);
    };
    // Reynolds-stress m-sum (radial_loop.f90:267-278)
    double part = 0.0;
    for (int i = threadIdx.x; i < nm; i += blockDim.x) {
        const int m = p.idx2m[i];
        const size_t q = (size_t)i * N + nr;
        const cd usf = m == 0 ? cscale(0.5 * p.ek * p.CorFac * p.ekpump[nr], p.up[q]) : p.us[q];
        part += cc22real(usf, p.om[q], m);
    }
    const double usom = block_reduce(part, red, 0);
    // Hermitian spectra of z = a + i b for (us, up), (om, T), (xi, 0)
    for (int t = threadIdx.x; t < 3 * n; t += blockDim.x) {
        const int tr = t / n, k = t - tr * n;
        if (tr == 2 && !p.l_chem) break;
        const cd *A = tr == 0 ? p.us : (tr == 1 ? p.om : p.xi);
        const cd *B = tr == 0 ? p.up : (tr == 1 ? (p.l_heat ? p.temp : nullptr) : nullptr);
        cd z = mk(0, 0);
        if (k == 0) {
            z = mk(A[nr].x, B ? B[nr].x : 0.0);
        } else if (k < nm) {
            const cd a = A[(size_t)k * N + nr], b = B ? B[(size_t)k * N + nr] : mk(0, 0);
            z = mk(a.x - b.y, a.y + b.x);
        } else if (n - k < nm) {
            const int kk = n - k;
            const cd a = cconj(A[(size_t)kk * N + nr]), b = B ? cconj(B[(size_t)kk * N + nr]) : mk(0, 0);
            z = mk(a.x - b.y, a.y + b.x);
        }
        (tr == 0 ? s0 : (tr == 1 ? s1 : s2))[k] = z;
    }
    __syncthreads();
    // one complex FFT at a time; the result lands in the column or in the
    // scratch, whose roles are then exchanged
    {
        cd *g = fft_stockham(s0, sc, n, 1, +1, p.radices, p.nrad, bwd);
        if (g != s0) { sc = s0; s0 = g; }
        __syncthreads();
        g = fft_stockham(s1, sc, n, 1, +1, p.radices, p.nrad, bwd);
        if (g != s1) { sc = s1; s1 = g; }
        if (p.l_chem) {
            __syncthreads();
            g = fft_stockham(s2, sc, n, 1, +1, p.radices, p.nrad, bwd);
            if (g != s2) { sc = s2; s2 = g; }
        }
        __syncthreads();
    }
    if (courant) {
        double vr = 0.0, vh = 0.0;
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            const cd z = s0[j];
            vr = fmax(vr, z.x * z.x);
            vh = fmax(vh, z.y * z.y);
        }
        vr = block_reduce(vr, red, 1);
        vh = block_reduce(vh, red, 1);
        if (threadIdx.x == 0) {
            const double cf = p.courfac;
            vr = cf * cf * vr;
            vh = cf * cf * vh;
            double dtrk = 1.0e10, dthk = 1.0e10;
            if (vr != 0.0) dtrk = fmin(dtrk, sqrt(p.delxr2[nr] / vr));
            if (vh != 0.0) dthk = fmin(dthk, sqrt(p.delxh2[nr] / vh));
            p.dtr[nr] = dtrk;
            p.dth[nr] = dthk;
        }
    }
    // nonlinear products in place (radial_loop.f90:316-325): (up X + i r us X)
    const double rr = p.r[nr];
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const cd z0 = s0[j], z1 = s1[j];
        const double us = z0.x, up = z0.y, om = z1.x, T = z1.y;
        const double X = p.l_chem ? s2[j].x : 0.0;
        s0[j] = mk(up * om, rr * us * om);
        s1[j] = mk(up * T, rr * us * T);
        if (p.l_chem) s2[j] = mk(up * X, rr * us * X);
    }
    __syncthreads();
    {
        cd *g = fft_stockham(s0, sc, n, 1, -1, p.radices, p.nrad, fwd);
        if (g != s0) { sc = s0; s0 = g; }
        if (p.l_heat) {
            __syncthreads();
            g = fft_stockham(s1, sc, n, 1, -1, p.radices, p.nrad, fwd);
            if (g != s1) { sc = s1; s1 = g; }
        }
        if (p.l_chem) {
            __syncthreads();
            g = fft_stockham(s2, sc, n, 1, -1, p.radices, p.nrad, fwd);
            if (g != s2) { sc = s2; s2 = g; }
        }
        __syncthreads();
    }
    // unpack, normalise and post-scale (radial_loop.f90:350-386)
    const double or1 = p.or1[nr];
    const double fbeta = (1.0 - p.tadvz) * p.beta[nr] * or1;
    for (int t = threadIdx.x; t < 3 * nm; t += blockDim.x) {
        const int tr = t / nm, i = t - tr * nm;
        if ((tr == 1 && !p.l_heat) || (tr == 2 && !p.l_chem)) continue;
        const cd *Z = tr == 0 ? s0 : (tr == 1 ? s1 : s2);
        const cd zk = Z[i], zc = cconj(Z[i == 0 ? 0 : n - i]);
        cd A = mk(0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y));
        cd B = mk(0.5 * (zk.y - zc.y), -0.5 * (zk.x - zc.x));
        A = cdivr(A, (double)n);
        B = cdivr(B, (double)n);
        const int m = p.idx2m[i];
        const size_t q = (size_t)i * N + nr;
        const cd cim = mk(0.0, -or1 * (double)m);
        if (tr == 0) {
            p.P_expl[is * F + q] = (i == 0) ? mk(-usom, 0.0) : cmul(cim, A);
            p.dVsOm[q] = B;
        } else if (tr == 1) {
            const cd v = cmul(cim, A);
            p.T_expl[is * F + q] = mk(v.x - fbeta * B.x, v.y - fbeta * B.y);
            p.dVsT[q] = B;
        } else {
            p.X_expl[is * F + q] = cmul(cim, A);
            p.dVsXi[q] = B;
        }
    }
}

// Long azimuthal grids without composition, where two columns of n_phi and
// the in-place transforms fit (fft_inplace_fits): the spectra by radius
// come from k_gather_r (contiguous rows), the pairs (u_s, u_phi) and
// (omega, T) travel through their in-place complex FFTs together, and LDS
// holds the two columns and a two-level twiddle table (64 + n_phi / 64
// entries).
__device__ __forceinline__ void radial_loop_inplace_body(const PzDev &p, int nr, int is, int courant)
{
    extern __shared__ double smem[];
    const int n = p.nphi, nm = p.nm, N = p.N;
    const size_t F = (size_t)nm * N;
    cd *s0 = (cd *)smem, *s1 = s0 + n, *t1 = s1 + n, *t2 = t1 + 64;
    const int n2 = (n + 63) >> 6;
    cd *om = t2 + n2;
    double *red = (double *)(om + PZ_FFT_OM_MAX), *msr = red + 16;
    const cd *Rus = p.rrow + (size_t)nr * nm, *Rup = Rus + F, *Rom = Rus + 2 * F, *RT = Rus + 3 * F;
    // exp(-2 pi i k / n) = t1[k mod 64] t2[k / 64]
    for (int k = threadIdx.x; k < 64 + n2; k += blockDim.x) t1[k] = p.twid[k < 64 ? k : 64 * (k - 64)];
    auto fwd = [=](int k) { return cmul(t1[k & 63], t2[k >> 6]); };
    auto bwd = [=](int k) {
        const cd w = cmul(t1[k & 63], t2[k >> 6]);
        return mk(w.x, -w.y);
    };
    // Hermitian spectra of z = a + i b for (us, up), (om, T), and the
    // Reynolds-stress m-sum (radial_loop.f90:267-278) from the same loads,
    // one partial sum per wave in msr (added at the unpack)
    const double ekc = 0.5 * p.ek * p.CorFac * p.ekpump[nr];
    double part = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        cd z0 = mk(0, 0), z1 = mk(0, 0);
        if (k < nm) {
            const cd us = Rus[k], up = Rup[k], o = Rom[k], T = p.l_heat ? RT[k] : mk(0, 0);
            const int m = p.idx2m[k];
            part += cc22real(m == 0 ? cscale(ekc, up) : us, o, m);
            if (k == 0) {
                z0 = mk(us.x, up.x);
                z1 = mk(o.x, T.x);
            } else {
                z0 = mk(us.x - up.y, us.y + up.x);
                z1 = mk(o.x - T.y, o.y + T.x);
            }
        } else if (n - k < nm) {
            const int kk = n - k;
            const cd us = cconj(Rus[kk]), up = cconj(Rup[kk]), o = cconj(Rom[kk]);
            const cd T = p.l_heat ? cconj(RT[kk]) : mk(0, 0);
            z0 = mk(us.x - up.y, us.y + up.x);
            z1 = mk(o.x - T.y, o.y + T.x);
        }
        s0[k] = z0;
        s1[k] = z1;
    }
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
    if ((threadIdx.x & 63) == 0) msr[threadIdx.x >> 6] = part;
    __syncthreads();
    fft_inplace(s0, n, 2, +1, p.radices, p.nrad, bwd, om);
    if (courant) {
        double vr = 0.0, vh = 0.0;
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            const cd z = s0[j];
            vr = fmax(vr, z.x * z.x);
            vh = fmax(vh, z.y * z.y);
        }
        vr = block_reduce(vr, red, 1);
        vh = block_reduce(vh, red, 1);
        if (threadIdx.x == 0) {
            const double cf = p.courfac;
            vr = cf * cf * vr;
            vh = cf * cf * vh;
            double dtrk = 1.0e10, dthk = 1.0e10;
            if (vr != 0.0) dtrk = fmin(dtrk, sqrt(p.delxr2[nr] / vr));
            if (vh != 0.0) dthk = fmin(dthk, sqrt(p.delxh2[nr] / vh));
            p.dtr[nr] = dtrk;
            p.dth[nr] = dthk;
        }
    }
    // nonlinear products in place (radial_loop.f90:316-325): (up X + i r us X)
    const double rr = p.r[nr];
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const cd z0 = s0[j], z1 = s1[j];
        const double us = z0.x, up = z0.y, om = z1.x, T = z1.y;
        s0[j] = mk(up * om, rr * us * om);
        s1[j] = mk(up * T, rr * us * T);
    }
    __syncthreads();
    fft_inplace(s0, n, p.l_heat ? 2 : 1, -1, p.radices, p.nrad, fwd, om);
    // unpack, normalise and post-scale (radial_loop.f90:350-386)
    const double or1 = p.or1[nr];
    const double fbeta = (1.0 - p.tadvz) * p.beta[nr] * or1;
    const int nout = p.l_heat ? 2 : 1;
    for (int t = threadIdx.x; t < nout * nm; t += blockDim.x) {
        const int tr = t >= nm, i = t - tr * nm;
        const cd *Z = tr == 0 ? s0 : s1;
        const cd zk = Z[i], zc = cconj(Z[i == 0 ? 0 : n - i]);
        cd A = mk(0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y));
        cd B = mk(0.5 * (zk.y - zc.y), -0.5 * (zk.x - zc.x));
        A = cdivr(A, (double)n);
        B = cdivr(B, (double)n);
        const int m = p.idx2m[i];
        const size_t q = (size_t)i * N + nr;
        const cd cim = mk(0.0, -or1 * (double)m);
        if (tr == 0) {
            if (i == 0) {
                double usom = msr[0];
                for (int w = 1; w < (int)((blockDim.x + 63) >> 6); w++) usom += msr[w];
                p.P_expl[is * F + q] = mk(-usom, 0.0);
            } else {
                p.P_expl[is * F + q] = cmul(cim, A);
            }
            p.dVsOm[q] = B;
        } else {
            const cd v = cmul(cim, A);
            p.T_expl[is * F + q] = mk(v.x - fbeta * B.x, v.y - fbeta * B.y);
            p.dVsT[q] = B;
        }
    }
}

// u_s, u_phi, omega (and T) from the mode-major layout [m][n_r] to rows by
// radius [f][n_r][n_m]: 16 x 32 tiles of all fields through LDS, coalesced
// on both sides, every load of a thread in flight before the one barrier
#define PZ_GATHER_TM 16
#define PZ_GATHER_TR 32
__global__ void __launch_bounds__(256) k_gather_r(const PzDev *__restrict__ pd)
{
    const PzDev &p = *pd;
    __shared__ cd tile[4][PZ_GATHER_TM][PZ_GATHER_TR + 1];
    const int N = p.N, nm = p.nm;
    const size_t F = (size_t)nm * N;
    const int m0 = blockIdx.x * PZ_GATHER_TM, r0 = blockIdx.y * PZ_GATHER_TR;
    const int nf = p.l_heat ? 4 : 3;
    constexpr int PER = PZ_GATHER_TM * PZ_GATHER_TR / 256; // elements a thread moves per field
    cd v[4][PER];
#pragma unroll
    for (int f = 0; f < 4; f++) {
        const cd *src = f == 0 ? p.us : f == 1 ? p.up : f == 2 ? p.om : p.temp;
#pragma unroll
        for (int e = 0; e < PER; e++) {
            const int t = threadIdx.x + e * 256, mm = t / PZ_GATHER_TR, rr = t % PZ_GATHER_TR;
            v[f][e] = (f < nf && m0 + mm < nm && r0 + rr < N) ? src[(size_t)(m0 + mm) * N + r0 + rr] : mk(0, 0);
        }
    }
#pragma unroll
    for (int f = 0; f < 4; f++)
#pragma unroll
        for (int e = 0; e < PER; e++) {
            const int t = threadIdx.x + e * 256;
            tile[f][t / PZ_GATHER_TR][t % PZ_GATHER_TR] = v[f][e];
        }
    __syncthreads();
    for (int f = 0; f < nf; f++) {
        cd *dst = p.rrow + f * F;
#pragma unroll
        for (int e = 0; e < PER; e++) {
            const int t = threadIdx.x + e * 256, rr = t / PZ_GATHER_TM, mm = t % PZ_GATHER_TM;
            if (m0 + mm < nm && r0 + rr < N) dst[(size_t)(r0 + rr) * nm + m0 + mm] = tile[f][mm][rr];
        }
    }
}

#define PZ_RADIAL_BIG_THREADS 1024
__global__ void __launch_bounds__(PZ_RADIAL_BIG_THREADS) k_radial_loop_big(const PzDev *__restrict__ pd, int is, int courant)
{
    radial_loop_big_body(*pd, blockIdx.x, is, courant);
}
// 512 threads within 128 registers: two workgroups (2 x 67 KB of LDS) share
// a CU, so the 257 radii of n_r = 257 run in one round on 256 CUs
#define PZ_RADIAL_INPLACE_THREADS 512
__global__ void __launch_bounds__(PZ_RADIAL_INPLACE_THREADS) __attribute__((amdgpu_waves_per_eu(4)))
k_radial_loop_inplace(const PzDev *__restrict__ pd, int is, int courant)
{
    radial_loop_inplace_body(*pd, blockIdx.x, is, courant);
}
// pz_fft / pz_ifft by the in-place transforms of the long grids, one
// radius per block (grid real [rad][n_phi], spectral [rad][nm])
__global__ void __launch_bounds__(PZ_RADIAL_INPLACE_THREADS)
k_phi_fft_inplace(const PzDev *__restrict__ pd, const double *grid, cd *spec)
{
    extern __shared__ double smem[];
    const PzDev &p = *pd;
    const int n = p.nphi, nm = p.nm, n2 = (n + 63) >> 6;
    cd *a = (cd *)smem, *t1 = a + 2 * n, *t2 = t1 + 64, *om = t2 + n2;
    for (int k = threadIdx.x; k < 64 + n2; k += blockDim.x) t1[k] = p.twid[k < 64 ? k : 64 * (k - 64)];
    const double *g = grid + (size_t)blockIdx.x * n;
    for (int j = threadIdx.x; j < n; j += blockDim.x) a[j] = mk(g[j], 0.0);
    __syncthreads();
    fft_inplace(a, n, 1, -1, p.radices, p.nrad, [=](int k) { return cmul(t1[k & 63], t2[k >> 6]); }, om);
    for (int i = threadIdx.x; i < nm; i += blockDim.x) spec[(size_t)blockIdx.x * nm + i] = cdivr(a[i], (double)n);
}
__global__ void __launch_bounds__(PZ_RADIAL_INPLACE_THREADS)
k_phi_ifft_inplace(const PzDev *__restrict__ pd, const cd *spec, double *grid)
{
    extern __shared__ double smem[];
    const PzDev &p = *pd;
    const int n = p.nphi, nm = p.nm, n2 = (n + 63) >> 6;
    cd *a = (cd *)smem, *t1 = a + 2 * n, *t2 = t1 + 64, *om = t2 + n2;
    for (int k = threadIdx.x; k < 64 + n2; k += blockDim.x) t1[k] = p.twid[k < 64 ? k : 64 * (k - 64)];
    const cd *s = spec + (size_t)blockIdx.x * nm;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        cd z = mk(0, 0);
        if (k == 0) z = mk(s[0].x, 0.0);
        else if (k < nm) z = s[k];
        else if (n - k < nm) z = cconj(s[n - k]);
        a[k] = z;
    }
    __syncthreads();
    fft_inplace(a, n, 1, +1, p.radices, p.nrad, [=](int k) {
        const cd w = cmul(t1[k & 63], t2[k >> 6]);
        return mk(w.x, -w.y);
    }, om);
    for (int j = threadIdx.x; j < n; j += blockDim.x) grid[(size_t)blockIdx.x * n + j] = a[j].x;
}

__global__ void __launch_bounds__(256) k_radial_loop(const PzDev *__restrict__ pd, int is, int courant)
{
    radial_loop_body(*pd, blockIdx.x, is, courant);
}
__global__ void __launch_bounds__(256) k_radial_loop_batch(PzBatch b, int is, int courant)
{
    const int2 e = b.map[blockIdx.x];
    radial_loop_body(*b.runs[e.x], e.y, is, courant);
}

// ---------------------------------------------------------------------------
// Courant decision and time-scheme weights on the device:
// dt_courant (courant.f90:78-167), set_dt_array and set_weights
// (multistep_schemes.f90:137-348; dirk_schemes.f90:275-771), lMat logic of
// step_time.f90:277-304.
// ---------------------------------------------------------------------------
// Weights of the multistep schemes for the dt history dt[0..3]
// (multistep_schemes.f90:137-305): rhs = sum wimp[o] old[o] +
// sum wimp_lin[o+1] impl[o] + sum wexp[o] expl[o], LHS weight wimp_lin[0].
__device__ void multistep_weights(int id, double alpha, const double *dt, double *wimp, double *wl, double *wexp)
{
    const double d0 = dt[0];
    switch (id) {
    case PZ_CNLF: {
        const double de = dt[0] / dt[1];
        wimp[0] = (1.0 - de) * (1.0 + de);
        wimp[1] = de * de;
        wl[0] = 0.5 * (1.0 + de) / de * d0;
        wl[1] = 0.5 * (1.0 + de) * (de - 1.0) / de * d0;
        wl[2] = 0.5 * (1.0 + de) * d0;
        wexp[0] = (1.0 + de) * d0;
        wexp[1] = 0.0;
        break;
    }
    case PZ_MODCNAB: {
        const double de = dt[0] / dt[1];
        wimp[0] = 1.0;
        wimp[1] = 0.0;
        wl[0] = (0.5 + 1.0 / de / 16.0) * d0;
        wl[1] = (7.0 / 16.0 - 1.0 / de / 16.0) * d0;
        wl[2] = 1.0 / 16.0 * d0;
        wexp[0] = (1.0 + 0.5 * de) * d0;
        wexp[1] = -0.5 * de * d0;
        break;
    }
    case PZ_SBDF2: {
        const double de = dt[0] / dt[1], q = 1.0 + 2.0 * de;
        wl[0] = (1.0 + de) / q * d0;
        wl[1] = 0.0;
        wimp[0] = (1.0 + de) * (1.0 + de) / q;
        wimp[1] = -de * de / q;
        wexp[0] = (1.0 + de) * (1.0 + de) * d0 / q;
        wexp[1] = -de * (1.0 + de) * d0 / q;
        break;
    }
    case PZ_SBDF3: {
        const double a = dt[1] / dt[0], b = dt[2] / dt[0];
        const double a0 = 1.0 + 1.0 / (1.0 + a) + 1.0 / (1.0 + a + b);
        const double a1 = (1.0 + a) * (1.0 + a + b) / (a * (a + b));
        const double a2 = -(1.0 + a + b) / (a * b * (1.0 + a));
        const double a3 = (1.0 + a) / (b * (a + b) * (1.0 + a + b));
        const double b0 = (1.0 + a) * (1.0 + a + b) / (a * (a + b));
        const double b1 = -(1.0 + a + b) / (a * b);
        const double b2 = (1.0 + a) / (b * (a + b));
        wl[0] = 1.0 / a0 * d0;
        wl[1] = 0.0;
        wimp[0] = a1 / a0; wimp[1] = a2 / a0; wimp[2] = a3 / a0;
        wexp[0] = b0 / a0 * d0; wexp[1] = b1 / a0 * d0; wexp[2] = b2 / a0 * d0;
        break;
    }
    case PZ_TVB33: {
        const double g = (2619.0 - sqrt(5995281.0)) / 345.0;
        const double th = (3525.0 * g * g - 3118.0 * g - 5084.0) / 5468.0;
        const double c = (2178.0 * th - 959.0 * g * g + 130.0 * g + 726.0) / 4096.0;
        const double a = dt[0] / dt[1], b = dt[1] / dt[2];
        const double a3 = -b * b * b * a * a * (3.0 * g * g * a + 2.0 * g * (1.0 - a) - 1.0) / (1.0 + b) /
                          (1.0 + b * (1.0 + a));
        const double a2 = -a * a * (g * a * b * (2.0 - 3.0 * g) + (1.0 - 2.0 * g) * (1.0 + b)) / (1.0 + a);
        const double a1 = -(b * (1.0 + g * a) * (1.0 + a * (3.0 * g - 2.0)) + a * (2.0 * g - 1.0) + 1.0) / (1.0 + b) - th;
        const double a0 = th + (1.0 + 2.0 * g * a + b * (1.0 + g * a) * (1.0 + 3.0 * g * a)) / (1.0 + a) /
                                   (1.0 + b * (1.0 + a));
        const double q = 1.0 + 0.5 * a + b * a * (3.0 + 2.0 * a) / (6.0 * (1.0 + b));
        const double b2 = b * b * a * (6.0 * g * (1.0 + g * a) + th * (3.0 + 2.0 * a)) / (6.0 * (1.0 + b));
        const double b1 = -g * a * (1.0 + b * (1.0 + g * a)) - th / 6.0 * a * (3.0 + b * (3.0 + 2.0 * a));
        const double b0 = (1.0 + g * a) * (1.0 + b * (1.0 + g * a)) / (1.0 + b) + th * q;
        const double c3 = th * b * b * a * (3.0 + 2.0 * a) / (6.0 * (1.0 + b)) - c;
        const double c2 = (c * (1.0 + b) * (1.0 + b * (1.0 + a)) - b * b * a * a * g * (1.0 - g)) / (b * b * (1.0 + a)) -
                          th * a * (3.0 + b * (3.0 + 2.0 * a)) / 6.0;
        const double c1 = (b * b * a * (1.0 - g) * (1.0 + g * a) - c * (1.0 + b * (1.0 + a))) / (b * b) / a + th * q;
        const double c0 = (b * b * a * g * (1.0 + g * a) + c * (1.0 + b)) / (b * b * a * (1.0 + a));
        wl[0] = c0 / a0 * d0; wl[1] = c1 / a0 * d0; wl[2] = c2 / a0 * d0; wl[3] = c3 / a0 * d0;
        wimp[0] = -a1 / a0; wimp[1] = -a2 / a0; wimp[2] = -a3 / a0;
        wexp[0] = b0 / a0 * d0; wexp[1] = b1 / a0 * d0; wexp[2] = b2 / a0 * d0;
        break;
    }
    case PZ_SBDF4: {
        const double a = dt[0] / dt[1], b = dt[1] / dt[2], e = dt[2] / dt[3];
        const double c1 = 1.0 + e * (1.0 + b), c2 = 1.0 + b * (1.0 + a), c3 = 1.0 + e * c2;
        const double a4 = (1.0 + a) / (1.0 + e) * c2 / c1 / c3 * (e * e * e * e * b * b * b * a * a);
        const double a3 = -b * b * b * a * a * (1.0 + a) / (1.0 + b) * c3 / c2;
        const double a2 = a * (a / (1.0 + a) + b * a * (c3 + e) / (1.0 + e));
        const double a1 = -1.0 - a * (1.0 + b * (1.0 + a) * (1.0 + e * c2 / c1) / (1.0 + b));
        const double a0 = 1.0 + a / (1.0 + a) + b * a / c2 + e * b * a / c3;
        const double b3 = -e * e * e * b * b * a * (1.0 + a) / (1.0 + e) * c2 / c1;
        const double b2 = b * b * a * (1.0 + a) / (1.0 + b) * c3;
        const double b1 = -c2 * c3 * a / (1.0 + e);
        const double b0 = b * (1.0 + a) / (1.0 + b) * ((1.0 + a) * (c3 + e) + (1.0 + e) / b) / c1;
        wl[0] = 1.0 / a0 * d0;
        wl[1] = 0.0;
        wimp[0] = -a1 / a0; wimp[1] = -a2 / a0; wimp[2] = -a3 / a0; wimp[3] = -a4 / a0;
        wexp[0] = b0 / a0 * d0; wexp[1] = b1 / a0 * d0; wexp[2] = b2 / a0 * d0; wexp[3] = b3 / a0 * d0;
        break;
    }
    default: // CNAB2
        wimp[0] = 1.0;
        wl[0] = alpha * d0;
        wl[1] = (1 - alpha) * d0;
        wexp[0] = (1.0 + 0.5 * dt[0] / dt[1]) * d0;
        wexp[1] = -0.5 * dt[0] * dt[0] / dt[1];
        break;
    }
}

__device__ void scheme_weights(const PzDev &p, PzScheme *s, int *lMatNext)
{
    double wimp_old = s->wimp_lin[0];
    if (p.family == 0) {
        multistep_weights(p.scheme, p.alpha, s->dt, s->wimp, s->wimp_lin, s->wexp);
        if (s->wimp_lin[0] != wimp_old) *lMatNext = 1;
    } else {
        s->wimp_lin[0] = s->dt[0] * p.wimp_lin0;
        for (int i = 0; i < PZ_MAX_STAGES; i++) {
            for (int j = 0; j < PZ_MAX_STAGES; j++) {
                s->bimp[i][j] = s->dt[0] * p.bimp0[i][j];
                s->bexp[i][j] = s->dt[0] * p.bexp0[i][j];
            }
            s->bass_imp[i] = s->dt[0] * p.bass_imp0[i];
            s->bass_exp[i] = s->dt[0] * p.bass_exp0[i];
        }
    }
}

__global__ void k_set_weights(const PzDev p __attribute__((grid_constant)))
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        PzScheme *s = p.ss;
        int dummy = 0;
        scheme_weights(p, s, &dummy);
        // dummy Courant limits before the first radial loop (step_time.f90:124-126)
        for (int i = 0; i < p.N; i++) { p.dtr[i] = 1.0e10; p.dth[i] = 1.0e10; }
    }
}

__device__ __forceinline__ void courant_body(const PzDev &p)
{
    __shared__ double red[8];
    PzScheme *s = p.ss;
    double dtr = 1.0e300, dth = 1.0e300;
    for (int i = threadIdx.x; i < p.N; i += blockDim.x) {
        dtr = fmin(dtr, p.dtr[i]);
        dth = fmin(dth, p.dth[i]);
    }
    dtr = -block_reduce(-dtr, red, 1);
    dth = -block_reduce(-dth, red, 1);
    if (threadIdx.x != 0) return;
    double dtMax = p.dtMax, dt = s->dt[0];
    double dt_r = fmin(dtr, 1000.0 * dtMax), dt_h = fmin(dth, 1000.0 * dtMax);
    double dt_rh = fmin(dt_r, dt_h);
    double dt_2 = fmin(0.5 * (1.0 / p.dt_fac + 1.0) * dt_rh, dtMax);
    int l_new_dt;
    double dt_new;
    if (dt > dtMax) { l_new_dt = 1; dt_new = dtMax; }
    else if (dt > dt_rh) { l_new_dt = 1; dt_new = dt_2; }
    else if (p.dt_fac * dt < dt_rh && dt < dtMax) { l_new_dt = 1; dt_new = dt_2; }
    else { l_new_dt = 0; dt_new = dt; }
    s->dt_r = dt_r;
    s->dt_h = dt_h;
    if (p.family == 0)
        for (int i = p.nexp - 1; i > 0; i--) s->dt[i] = s->dt[i - 1];
    s->dt[0] = dt_new;
    if (dt_new < p.dtMin) s->error = 1;
    int lMatNext = s->lMatNext;
    scheme_weights(p, s, &lMatNext);
    s->time += s->dt[0];
    if (p.bridge && s->n_steps + 1 <= p.nold - 1) {
        // bridge_with_cnab2 (multistep_schemes.f90:500-524): this step uses
        // the CNAB2 weights, the higher-order ones are zero
        double w[8] = {}, wl[8] = {}, we[8] = {};
        multistep_weights(PZ_CNAB2, p.alpha, s->dt, w, wl, we);
        for (int i = 0; i < 8; i++) {
            s->wimp[i] = i < 1 ? w[i] : 0.0;
            s->wimp_lin[i] = i < 2 ? wl[i] : 0.0;
            s->wexp[i] = i < 2 ? we[i] : 0.0;
        }
    }
    if (s->ab1 && s->n_steps == 0 && p.family == 0) {
        // start_with_ab1 (multistep_schemes.f90:526-534): the first step
        // after a restart that changed ek or the radial method
        s->wexp[0] = s->dt[0];
        for (int i = 1; i < 8; i++) s->wexp[i] = 0.0;
        s->ab1 = 0;
    }
    int lMat = lMatNext;
    if (l_new_dt || lMat) lMat = 1;
    s->lMat = lMat;
    s->lMatNext = 0;
    s->l_new_dt = l_new_dt;
    s->n_steps += 1;
}

__global__ void __launch_bounds__(256) k_courant(const PzDev *__restrict__ pd) { courant_body(*pd); }
__global__ void __launch_bounds__(256) k_courant_batch(PzBatch b) { courant_body(*b.runs[b.map[blockIdx.x].x]); }

// ---------------------------------------------------------------------------
// Wave-level banded solve (dgbtrs 'N' semantics: row interchanges with L,
// then the banded upper solve of dtbsv) on one right-hand side in LDS.
// Lane l of the wave holds element j0 + l of a 64-wide window; each step
// broadcasts the pivot row element with v_readlane and updates kl (or
// kl+ku) lanes, so no LDS traffic sits on the recurrence.
// ---------------------------------------------------------------------------
template <class M>
__device__ void wave_gbtrs(int n, int kl, int ku, const M *ab, int ld, const int *ipiv, cd *b)
{
    const int lane = threadIdx.x & 63;
    const int kd = kl + ku;
    int j = 0;
    while (j < n - 1) {
        const int j0 = j;
        cd v = (j0 + lane < n) ? b[j0 + lane] : mk(0, 0);
        const int jend = min(n - 1, j0 + 64 - kl);
        for (; j < jend; j++) {
            const int h = j - j0;
            const int lm = min(kl, n - 1 - j);
            const int pv = ipiv[j] - j;
            if (pv != 0) {
                cd a = rlc(v, h), c = rlc(v, h + pv);
                if (lane == h) v = c;
                else if (lane == h + pv) v = a;
            }
            const cd bj = rlc(v, h);
            const int d = lane - h;
            if (d >= 1 && d <= lm) v = csub(v, mmul(ab[kd + d + (size_t)j * ld], bj));
        }
        if (j0 + lane < n) b[j0 + lane] = v;
        __builtin_amdgcn_wave_barrier();
    }
    const int k = kd;
    j = n - 1;
    while (j >= 0) {
        const int j0 = j;
        cd v = (j0 - lane >= 0) ? b[j0 - lane] : mk(0, 0);
        const int jend = max(-1, j0 - (64 - k));
        for (; j > jend; j--) {
            const int h = j0 - j;
            cd xj = rlc(v, h);
            if (mnz(xj)) xj = mdiv(xj, ab[kd + (size_t)j * ld]);
            if (lane == h) v = xj;
            const int d = lane - h;
            if (d >= 1 && d <= min(k, j)) v = csub(v, mmul(ab[kd - d + (size_t)j * ld], xj));
        }
        if (j0 - lane >= 0) b[j0 - lane] = v;
        __builtin_amdgcn_wave_barrier();
    }
}

// small dense solve (dgetrs 'N'), one thread; a column-major nt x nt LU
template <class M>
__device__ void small_getrs(int nt, const M *a, const int *piv, cd *b)
{
    for (int i = 0; i < nt; i++)
        if (piv[i] != i) { cd t = b[i]; b[i] = b[piv[i]]; b[piv[i]] = t; }
    for (int jj = 0; jj < nt; jj++)
        for (int i = jj + 1; i < nt; i++) b[i] = csub(b[i], mmul(a[i + jj * nt], b[jj]));
    for (int jj = nt - 1; jj >= 0; jj--) {
        b[jj] = mdiv(b[jj], a[jj + jj * nt]);
        for (int i = 0; i < jj; i++) b[i] = csub(b[i], mmul(a[i + jj * nt], b[jj]));
    }
}


template <class M> __device__ __forceinline__ M rlm(M v, int lane);
template <> __device__ __forceinline__ double rlm<double>(double v, int lane) { return rl(v, lane); }
template <> __device__ __forceinline__ cd rlm<cd>(cd v, int lane) { return rlc(v, lane); }

// Banded solve with the factors in LDS (same semantics as wave_gbtrs).
// Off the recurrence: the next step's band coefficient is fetched one step
// ahead, the pivot indices and the reciprocal diagonal of U sit in lane
// registers and are broadcast with v_readlane.
template <class M, int KL, int KU>
__device__ void wave_gbtrs_fast(int n, const M *ab, int ld, const int *ipiv, const M *dinv, cd *b)
{
    const int lane = threadIdx.x & 63;
    constexpr int kd = KL + KU, K = KL + KU;
    int j = 0;
    while (j < n - 1) {
        const int j0 = j;
        cd v = (j0 + lane < n) ? b[j0 + lane] : mk(0, 0);
        const int preg = (j0 + lane < n) ? ipiv[j0 + lane] : 0;
        const int jend = min(n - 1, j0 + 64 - KL);
        const int base = kd + lane + j0 * ld;
        M cn = mzero<M>();
        if (lane >= 1 && lane <= KL && j0 + lane < n) cn = ab[base];
        for (int h = 0; j < jend; j++, h++) {
            const M c = cn;
            {
                const int dn = lane - h - 1;
                cn = (dn >= 1 && dn <= KL && j + 1 + dn < n) ? ab[base + (h + 1) * (ld - 1)] : mzero<M>();
            }
            const int pv = __builtin_amdgcn_readlane(preg, h) - j;
            if (pv != 0) {
                cd a = rlc(v, h), e = rlc(v, h + pv);
                if (lane == h) v = e;
                else if (lane == h + pv) v = a;
            }
            const cd bj = rlc(v, h);
            const int d = lane - h;
            if (d >= 1 && d <= KL && j + d < n) v = csub(v, mmul(c, bj));
        }
        if (j0 + lane < n) b[j0 + lane] = v;
        __builtin_amdgcn_wave_barrier();
    }
    j = n - 1;
    while (j >= 0) {
        const int j0 = j;
        cd v = (j0 - lane >= 0) ? b[j0 - lane] : mk(0, 0);
        const M dreg = (j0 - lane >= 0) ? dinv[j0 - lane] : mzero<M>();
        const int jend = max(-1, j0 - (64 - K));
        const int base = kd - lane + j0 * ld;
        M cn = mzero<M>();
        if (lane >= 1 && lane <= K && j0 - lane >= 0) cn = ab[base];
        for (int h = 0; j > jend; j--, h++) {
            const M c = cn;
            {
                const int dn = lane - h - 1;
                cn = (dn >= 1 && dn <= K && j - 1 - dn >= 0) ? ab[base - (h + 1) * (ld - 1)] : mzero<M>();
            }
            const cd xj = mmul(rlm<M>(dreg, h), rlc(v, h));
            if (lane == h) v = xj;
            const int d = lane - h;
            if (d >= 1 && d <= K && j - d >= 0) v = csub(v, mmul(c, xj));
        }
        if (j0 - lane >= 0) b[j0 - lane] = v;
        __builtin_amdgcn_wave_barrier();
    }
}

// bordered solve with all factors in LDS
template <class M, int KL, int KU>
__device__ void wave_bordered_fast(int N, int nt, const M *A1, const int *p1, const M *A2, const M *A3,
                                   const M *A4, const int *p4, const M *dinv, cd *rhs)
{
    const int lane = threadIdx.x & 63;
    const int nb = N - nt, ld = 2 * KL + KU + 1;
    cd *r2 = rhs + nt;
    wave_gbtrs_fast<M, KL, KU>(nb, A4, ld, p4, dinv, r2);
    for (int i = 0; i < nt; i++) {
        double sx = 0.0, sy = 0.0;
        for (int c = lane; c < nb; c += 64) {
            cd t = mmul(A2[i + (size_t)c * nt], r2[c]);
            sx += t.x;
            sy += t.y;
        }
        sx = wave_sum(sx);
        sy = wave_sum(sy);
        if (lane == 0) rhs[i] = mk(rhs[i].x - sx, rhs[i].y - sy);
    }
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) small_getrs<M>(nt, A1, p1, rhs);
    __builtin_amdgcn_wave_barrier();
    cd y[4];
    for (int c = 0; c < nt; c++) y[c] = rhs[c];
    for (int i = lane; i < nb; i += 64) {
        cd v = r2[i];
        for (int c = 0; c < nt; c++) v = csub(v, mmul(A3[i + (size_t)c * nb], y[c]));
        r2[i] = v;
    }
    __builtin_amdgcn_wave_barrier();
}

// bordered solve (algebra_lapack.f90:332-416) by one wave; rhs in LDS
template <class M>
__device__ void wave_bordered_solve(int N, int nt, int kl, int ku, int ld, const M *A1, const int *p1,
                                    const M *A2, const M *A3, const M *A4, const int *p4, cd *rhs)
{
    const int lane = threadIdx.x & 63;
    const int nb = N - nt;
    cd *r2 = rhs + nt;
    wave_gbtrs<M>(nb, kl, ku, A4, ld, p4, r2);
    // rhs1 <- rhs1 - A2 r2
    for (int i = 0; i < nt; i++) {
        double sx = 0.0, sy = 0.0;
        for (int c = lane; c < nb; c += 64) {
            cd t = mmul(A2[i + (size_t)c * nt], r2[c]);
            sx += t.x;
            sy += t.y;
        }
        sx = wave_sum(sx);
        sy = wave_sum(sy);
        if (lane == 0) rhs[i] = mk(rhs[i].x - sx, rhs[i].y - sy);
    }
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) small_getrs<M>(nt, A1, p1, rhs);
    __builtin_amdgcn_wave_barrier();
    cd y[4];
    for (int c = 0; c < nt; c++) y[c] = rhs[c];
    for (int i = lane; i < nb; i += 64) {
        cd v = r2[i];
        for (int c = 0; c < nt; c++) v = csub(v, mmul(A3[i + (size_t)c * nb], y[c]));
        r2[i] = v;
    }
    __builtin_amdgcn_wave_barrier();
}

// ---------------------------------------------------------------------------
// Chebyshev transforms on LDS vectors, whole workgroup (dct_fftw.f90)
// costf (DCT-I with cheb_fac) of nf vectors: out[f][k]
__device__ void dct_many(const PzDev &p, VecSet<const cd> in, VecSet<cd> out, int nf, int dealias,
                         const double *cosT = nullptr)
{
    const int N = p.N, L = p.L;
    if (!cosT) cosT = p.cosT;
    for (int t = threadIdx.x; t < nf * N; t += blockDim.x) {
        int f = t / N, k = t - f * N;
        const cd *x = in[f];
        double sx = 0.0, sy = 0.0;
        int q = k;
#pragma unroll 4
        for (int j = 1; j < N - 1; j++, q += k) {
            if (q >= L) q -= L;
            double c = cosT[q];
            sx += x[j].x * c;
            sy += x[j].y * c;
        }
        double sg = (k & 1) ? -1.0 : 1.0;
        double ax = x[0].x + sg * x[N - 1].x + 2.0 * sx;
        double ay = x[0].y + sg * x[N - 1].y + 2.0 * sy;
        cd r = mk(p.cheb_fac * ax, p.cheb_fac * ay);
        if (dealias && k >= p.ncheb) r = mk(0, 0);
        out[f][k] = r;
    }
}

// Radial derivatives from Chebyshev coefficients w (get_dr_fft /
// get_ddr_fft with l_dct_in = .false., dct_fftw.f90), split in three parts:
// the boundary points x = +-1 (closed k^2 sums), the bulk points from the
// sums V = sum_{k=1}^{N-2} k w_k sin(pi k j/(N-1)) and U = sum k^2 w_k
// cos(pi k j/(N-1)), and the map to r by drx/ddrx.
__device__ __forceinline__ void dr_boundary(const PzDev &p, const cd *w, int j, bool second, cd &d1, cd &d2)
{
    const int N = p.N;
    const double cf = p.cheb_fac;
    double ax = 0, ay = 0, bx = 0, by = 0;
    for (int k = 1; k <= N - 1; k++) {
        double wx = w[k].x, wy = w[k].y;
        if (k == N - 1) { wx *= 0.5; wy *= 0.5; }
        const double s1 = (j == 0) ? 1.0 : (((k + 1) & 1) ? -1.0 : 1.0);
        const double k2 = (double)k * k;
        ax += s1 * k2 * wx;
        ay += s1 * k2 * wy;
        if (second && k >= 2) {
            const double s2 = (j == 0) ? 1.0 : ((k & 1) ? -1.0 : 1.0);
            const double kk = k2 * (k2 - 1.0);
            bx += s2 * kk * wx;
            by += s2 * kk * wy;
        }
    }
    // the coefficient vector enters divided by cheb_fac
    ax /= cf; ay /= cf; bx /= cf; by /= cf;
    d1 = mk(ax / (double)(N - 1), ay / (double)(N - 1));
    d2 = second ? mk((1.0 / 3.0) * bx / (double)(N - 1), (1.0 / 3.0) * by / (double)(N - 1)) : mk(0, 0);
}

// bulk point j from V = vx + i vy and U = ux + i uy (sums over the raw
// coefficients); drb holds 1/(sqrt(1-x^2) nrm), x/((1-x^2)^1.5 nrm) and
// 1/((1-x^2) nrm) with nrm = 2 N - 2
__device__ __forceinline__ void dr_bulk(const PzDev &p, const RadView &R, cd wl, int j, bool second, double vx, double vy,
                                        double ux, double uy, cd &d1, cd &d2)
{
    const int N = p.N;
    const double icf = p.icheb_fac;
    vx *= icf; vy *= icf; ux *= icf; uy *= icf;
    const double V1x = -2.0 * vx, V1y = -2.0 * vy;
    const double g1 = R.drb(j), g2 = R.drb(N + j), g3 = R.drb(2 * N + j);
    d1 = mk(-V1x * g1, -V1y * g1);
    d2 = mk(0, 0);
    if (second) {
        const double wlx = wl.x * icf, wly = wl.y * icf;
        const double sgn = (j & 1) ? -1.0 : 1.0;
        const double n1 = (double)(N - 1) * (double)(N - 1);
        const double W2x = -(2.0 * ux + n1 * wlx * sgn), W2y = -(2.0 * uy + n1 * wly * sgn);
        d2 = mk(-V1x * g2 + W2x * g3, -V1y * g2 + W2y * g3);
    }
}

__device__ __forceinline__ void dr_store(const PzDev &p, const RadView &R, int j, cd d1, cd d2, bool second, cd *df, cd *ddf)
{
    const double a = R.drx(j);
    if (second) {
        const double b = R.ddrx(j);
        lds_store(ddf, j, mk(b * d1.x + a * a * d2.x, b * d1.y + a * a * d2.y));
    }
    lds_store(df, j, mk(a * d1.x, a * d1.y));
}

// first (and optionally second) derivative of nf vectors, direct sums
__device__ void dr_many(const PzDev &p, const RadView &R, VecSet<const cd> c, VecSet<cd> df, VecSet<cd> ddf, int nf,
                        const double *cosT = nullptr, const double *sinT = nullptr)
{
    const int N = p.N, L = p.L;
    if (!cosT) cosT = p.cosT;
    if (!sinT) sinT = p.sinT;
    for (int t = threadIdx.x; t < nf * N; t += blockDim.x) {
        const int f = t / N, j = t - f * N;
        const cd *w = c[f];
        cd d1, d2;
        if (j == 0 || j == N - 1) {
            dr_boundary(p, w, j, ddf.v0 != nullptr, d1, d2);
        } else {
            double vx = 0, vy = 0, ux = 0, uy = 0;
            int q = j;
#pragma unroll 4
            for (int k = 1; k < N - 1; k++, q += j) {
                if (q >= L) q -= L;
                const double wx = w[k].x, wy = w[k].y;
                const double s = sinT[q], co = cosT[q];
                vx += (double)k * wx * s;
                vy += (double)k * wy * s;
                if (ddf.v0) {
                    const double k2 = (double)k * k;
                    ux += k2 * wx * co;
                    uy += k2 * wy * co;
                }
            }
            dr_bulk(p, R, w[N - 1], j, ddf.v0 != nullptr, vx, vy, ux, uy, d1, d2);
        }
        dr_store(p, R, j, d1, d2, ddf.v0 != nullptr, df[f], ddf.v0 ? ddf[f] : df[f]);
    }
}

// costf of nf <= 3 vectors through one FFT of length L = 2 N - 2 of the
// even extension y_j = x_j, y_{L-j} = x_j: Y_k = x_0 + (-1)^k x_{N-1}
// + 2 sum_{j=1}^{N-2} x_j cos(pi j k/(N-1)), the DCT-I of dct_fftw.f90.
// buf: LDS scratch of 2 nf L complex values; no trailing barrier.
__device__ __forceinline__ void dct_fft(const PzDev &p, VecSet<const cd> in, VecSet<cd> out, int nf, int dealias,
                                        const double *cosT, const double *sinT, cd *buf)
{
    if (p.nradL == 0) {
        dct_many(p, in, out, nf, dealias, cosT);
        return;
    }
    const int N = p.N, L = p.L;
    if (const int P = wave_fft_p(p)) {
        // even extension in, costf out, straight from and to the vectors
        // the vector pointers as plain values (a VecSet inside the closures
        // would be kept in scratch memory)
        const double cf = p.cheb_fac;
        const int nc = dealias ? p.ncheb : N;
        const cd *i0 = in.v0, *i1 = in.v1, *i2 = in.v2, *i3 = in.v3;
        cd *o0 = out.v0, *o1 = out.v1, *o2 = out.v2, *o3 = out.v3;
        fft_wave_p(
            P, nf,
            [=](int f, int j) __attribute__((always_inline)) {
                const cd *x = f == 0 ? i0 : f == 1 ? i1 : f == 2 ? i2 : i3;
                return lds_load(x, j < N ? j : L - j);
            },
            [=](int f, int k, cd z) __attribute__((always_inline)) {
                cd *y = f == 0 ? o0 : f == 1 ? o1 : f == 2 ? o2 : o3;
                if (k < N) lds_store(y, k, k >= nc ? mk(0.0, 0.0) : mk(cf * z.x, cf * z.y));
            },
            [=](int q) __attribute__((always_inline)) { return mk(cosT[q], -sinT[q]); });
        return;
    }
    cd *A = buf, *B = buf + (size_t)nf * L;
    for (int t = threadIdx.x; t < nf * L; t += blockDim.x) {
        const int f = t / L, j = t - f * L;
        A[t] = lds_load(in[f], j < N ? j : L - j);
    }
    __syncthreads();
    const cd *Z = fft_cheb(p, A, B, nf, -1, cosT, sinT);
    const double cf = p.cheb_fac;
    for (int t = threadIdx.x; t < nf * N; t += blockDim.x) {
        const int f = t / N, k = t - f * N;
        const cd z = Z[(size_t)f * L + k];
        lds_store(out[f], k, (dealias && k >= p.ncheb) ? mk(0.0, 0.0) : mk(cf * z.x, cf * z.y));
    }
}

// derivatives of nf <= 3 vectors through FFTs of length L: the odd
// extension of k w_k gives -2i V, the even extension of k^2 w_k (zero at
// k = 0, N-1) gives 2U.  The boundary points x = +-1 are k^2-weighted sums,
// one half-wave reduction per (vector, side).  With dct_out (nf = 1) the
// same FFT pass also returns costf(c) (even extension of c itself).
// buf: LDS scratch of 2 ntr L + 4 nf complex values (ntr = nf, + nf with
// the second derivative, + 1 with dct_out); no trailing barrier.
__device__ __forceinline__ void dr_fft(const PzDev &p, const RadView &R, VecSet<const cd> c, VecSet<cd> df,
                                       VecSet<cd> ddf, int nf, const double *cosT, const double *sinT, cd *buf,
                                       cd *dct_out = nullptr)
{
    if (p.nradL == 0) {
        dr_many(p, R, c, df, ddf, nf, cosT, sinT);
        if (dct_out) dct_many(p, VecSet<const cd>(c[0]), VecSet<cd>(dct_out), 1, 0, cosT);
        return;
    }
    const bool second = ddf.v0 != nullptr, with_dct = dct_out != nullptr;
    const int N = p.N, L = p.L, nder = second ? 2 * nf : nf, ntr = nder + (with_dct ? 1 : 0);
    cd *A = buf, *B = buf + (size_t)ntr * L, *bnd = buf + (size_t)2 * ntr * L;
    // input j of transform tr: the odd extension of k w_k (tr < nf), the
    // even extension of k^2 w_k (nf <= tr < nder), c itself (dct_out)
    const cd *c0 = c.v0, *c1 = c.v1, *c2 = c.v2, *c3 = c.v3;
    auto in_at = [=](int tr, int j) __attribute__((always_inline)) {
        const int f = tr < nf ? tr : (tr < nder ? tr - nf : 0);
        const int k = j < N ? j : L - j;
        cd y = mk(0.0, 0.0);
        if (tr >= nder) {
            y = lds_load(c0, k);
        } else if (k != 0 && k != N - 1) {
            const cd w = lds_load(f == 0 ? c0 : f == 1 ? c1 : f == 2 ? c2 : c3, k);
            const double sc = tr < nf ? (j < N ? (double)k : -(double)k) : (double)k * (double)k;
            y = mk(sc * w.x, sc * w.y);
        }
        return y;
    };
    const int P = wave_fft_p(p);
    if (P) {
        // transforms straight into A (natural order), one wave each
        fft_wave_p(
            P, ntr, in_at, [=](int tr, int k, cd z) __attribute__((always_inline)) { A[(size_t)tr * L + k] = z; },
            [=](int q) __attribute__((always_inline)) { return mk(cosT[q], -sinT[q]); });
    } else {
        for (int t = threadIdx.x; t < ntr * L; t += blockDim.x) {
            const int tr = t / L, j = t - tr * L;
            A[t] = in_at(tr, j);
        }
    }
    {
        // half-waves counted from the last wave, away from the waves that
        // run the register FFTs
        const int seg = (int)((blockDim.x - 1 - threadIdx.x) >> 5), hl = threadIdx.x & 31;
        if (seg < 2 * nf) {
            const int f = seg >> 1, side = seg & 1; // side 0: x = 1 (j = 0), 1: x = -1 (j = N-1)
            const cd *w = c[f];
            double ax = 0, ay = 0, bx = 0, by = 0;
            for (int k = 1 + hl; k <= N - 1; k += 32) {
                const cd wk = lds_load(w, k);
                double wx = wk.x, wy = wk.y;
                if (k == N - 1) { wx *= 0.5; wy *= 0.5; }
                const double k2 = (double)k * k;
                const double s1 = side == 0 ? 1.0 : ((k & 1) ? 1.0 : -1.0);
                ax += s1 * k2 * wx;
                ay += s1 * k2 * wy;
                if (second && k >= 2) {
                    const double s2 = side == 0 ? 1.0 : ((k & 1) ? -1.0 : 1.0);
                    const double kk = k2 * (k2 - 1.0);
                    bx += s2 * kk * wx;
                    by += s2 * kk * wy;
                }
            }
            for (int o = 16; o > 0; o >>= 1) {
                ax += __shfl_xor(ax, o, 32);
                ay += __shfl_xor(ay, o, 32);
                bx += __shfl_xor(bx, o, 32);
                by += __shfl_xor(by, o, 32);
            }
            if (hl == 0) {
                const double sc = p.icheb_fac / (double)(N - 1);
                bnd[2 * seg] = mk(ax * sc, ay * sc);
                bnd[2 * seg + 1] = mk((1.0 / 3.0) * bx * sc, (1.0 / 3.0) * by * sc);
            }
        }
    }
    __syncthreads();
    const cd *Z = P ? A : fft_cheb(p, A, B, ntr, -1, cosT, sinT);
    for (int t = threadIdx.x; t < nf * N; t += blockDim.x) {
        const int f = t / N, j = t - f * N;
        cd d1, d2;
        if (j == 0 || j == N - 1) {
            const int seg = 2 * f + (j == 0 ? 0 : 1);
            d1 = bnd[2 * seg];
            d2 = second ? bnd[2 * seg + 1] : mk(0.0, 0.0);
        } else {
            const cd zv = Z[(size_t)f * L + j];
            const cd zu = second ? Z[(size_t)(nf + f) * L + j] : mk(0.0, 0.0);
            dr_bulk(p, R, lds_load(c[f], N - 1), j, second, -0.5 * zv.y, 0.5 * zv.x, 0.5 * zu.x, 0.5 * zu.y, d1, d2);
        }
        dr_store(p, R, j, d1, d2, second, df[f], second ? ddf[f] : df[f]);
    }
    if (with_dct) {
        const double cf = p.cheb_fac;
        for (int k = threadIdx.x; k < N; k += blockDim.x) {
            const cd z = Z[(size_t)nder * L + k];
            lds_store(dct_out, k, mk(cf * z.x, cf * z.y));
        }
    }
}

// y = band * x for a row-major band [n][2K+1]; rows < nzero become zero.
// Four lanes share a row: lane s takes the diagonals s, s+4, ... (loads
// issued together), and two xor-shuffles combine the partial sums in a
// fixed order.
template <int K>
__device__ __forceinline__ void bandmv(const PzDev &p, const double *band, const cd *x, cd *y, int nzero)
{
    constexpr int W = 2 * K + 1, G = 4, PER = (W + G - 1) / G;
    const int N = p.N, sub = threadIdx.x & (G - 1);
    for (int i = threadIdx.x / G; i < N; i += blockDim.x / G) {
        double sx = 0.0, sy = 0.0;
        if (i >= nzero) {
            double b[PER];
#pragma unroll
            for (int u = 0; u < PER; u++) {
                const int d = sub + u * G, jj = i - K + d;
                b[u] = (d < W && jj >= 0 && jj < N) ? gld(band, (size_t)i * W + d) : 0.0;
            }
#pragma unroll
            for (int u = 0; u < PER; u++) {
                const int d = sub + u * G, jj = i - K + d;
                if (d < W && jj >= 0 && jj < N) {
                    const cd xv = lds_load(x, jj);
                    sx += b[u] * xv.x;
                    sy += b[u] * xv.y;
                }
            }
        }
        sx += __shfl_xor(sx, 1, G);
        sy += __shfl_xor(sy, 1, G);
        sx += __shfl_xor(sx, 2, G);
        sy += __shfl_xor(sy, 2, G);
        if (sub == 0) y[i] = mk(sx, sy);
    }
}
template <int K>
__device__ __forceinline__ void bandmv_c(const PzDev &p, const cd *band, const cd *x, cd *y, int nzero)
{
    constexpr int W = 2 * K + 1, G = 4, PER = (W + G - 1) / G;
    const int N = p.N, sub = threadIdx.x & (G - 1);
    for (int i = threadIdx.x / G; i < N; i += blockDim.x / G) {
        cd s = mk(0, 0);
        if (i >= nzero) {
            cd b[PER];
#pragma unroll
            for (int u = 0; u < PER; u++) {
                const int d = sub + u * G, jj = i - K + d;
                b[u] = (d < W && jj >= 0 && jj < N) ? gld(band, (size_t)i * W + d) : mk(0.0, 0.0);
            }
#pragma unroll
            for (int u = 0; u < PER; u++) {
                const int d = sub + u * G, jj = i - K + d;
                if (d < W && jj >= 0 && jj < N) s = cadd(s, cmul(b[u], lds_load(x, jj)));
            }
        }
        s.x += __shfl_xor(s.x, 1, G);
        s.y += __shfl_xor(s.y, 1, G);
        s.x += __shfl_xor(s.x, 2, G);
        s.y += __shfl_xor(s.y, 2, G);
        if (sub == 0) y[i] = s;
    }
}

// ---------------------------------------------------------------------------
// Device assembly + factorisation of the bordered LHS matrices
// (get_lhs_mat_tau: update_temp_integ.f90:631-776,
//  update_psi_integ_smat.f90:1286-1581; prepare_bordered_mat:
//  algebra_lapack.f90:216-286).  One wave per (field, m).
// ---------------------------------------------------------------------------
template <class M>
__device__ void wave_gbtrf(int n, int kl, int ku, M *ab, int *ipiv, int *info)
{
    const int lane = threadIdx.x & 63;
    const int ld = 2 * kl + ku + 1, kv = ku + kl;
    __shared__ int s_jp, s_ju;
    if (lane == 0) { *info = 0; s_ju = 0; }
    for (int jj = ku + 1; jj < min(kv, n); jj++)
        for (int i = kv - jj + lane; i < kl; i += 64) ab[i + (size_t)jj * ld] = mzero<M>();
    __syncthreads();
    for (int j = 0; j < n; j++) {
        if (j + kv < n)
            for (int i = lane; i < kl; i += 64) ab[i + (size_t)(j + kv) * ld] = mzero<M>();
        const int km = min(kl, n - 1 - j);
        __syncthreads();
        if (lane == 0) {
            int jp = 0;
            double mx = mabs1(ab[kv + (size_t)j * ld]);
            for (int i = 1; i <= km; i++) {
                double a = mabs1(ab[kv + i + (size_t)j * ld]);
                if (a > mx) { mx = a; jp = i; }
            }
            s_jp = jp;
            ipiv[j] = jp + j;
            if (mnz(ab[kv + jp + (size_t)j * ld])) {
                int t = min(j + ku + jp, n - 1);
                if (t > s_ju) s_ju = t;
            } else if (*info == 0) {
                *info = j + 1;
            }
        }
        __syncthreads();
        const int jp = s_jp, ju = s_ju;
        if (!mnz(ab[kv + jp + (size_t)j * ld])) continue;
        if (jp != 0)
            for (int c = lane; c <= ju - j; c += 64) {
                M t = ab[kv + jp - c + (size_t)(j + c) * ld];
                ab[kv + jp - c + (size_t)(j + c) * ld] = ab[kv - c + (size_t)(j + c) * ld];
                ab[kv - c + (size_t)(j + c) * ld] = t;
            }
        __syncthreads();
        if (km > 0) {
            M rp = mrecip(ab[kv + (size_t)j * ld]);
            for (int i = 1 + lane; i <= km; i += 64) ab[kv + i + (size_t)j * ld] = mmm(ab[kv + i + (size_t)j * ld], rp);
            __syncthreads();
            const int nc = ju - j;
            for (int t = lane; t < km * nc; t += 64) {
                int c = 1 + t / km, i = 1 + t % km;
                M temp = mneg(ab[kv - c + (size_t)(j + c) * ld]);
                ab[kv + i - c + (size_t)(j + c) * ld] =
                    madd(ab[kv + i - c + (size_t)(j + c) * ld], mmm(ab[kv + i + (size_t)j * ld], temp));
            }
        }
        __syncthreads();
    }
}

template <class M> __device__ M mk_elem(double re, double im);
template <> __device__ double mk_elem<double>(double re, double) { return re; }
template <> __device__ cd mk_elem<cd>(double re, double im) { return mk(re, im); }

// Assembles and factorises one bordered matrix from raw stencils.
//   st = S0 - c1*S1 (+ i c2*S2), rows idx = n_r + nt - 1
// tau rows taken from tauv[nt][N]; result written to global storage.
template <class M>
__device__ void build_bordered(const PzDev &p, int nt, int kl, int ku, int K, const double *S0,
                               const double *S1, const double *S2, double c1, double c2,
                               const double *tauv, M *A1g, M *A2g, M *A3g, M *A4g, int *p1g,
                               int *p4g, M *dinvg, double *facg, int *infog, char *lds)
{
    const int N = p.N, nb = N - nt, ld = 2 * kl + ku + 1, W = 2 * K + 1;
    const int lane = threadIdx.x & 63;
    // LDS carve-up, every array 16-byte aligned
    size_t o = 0;
    auto take = [&](size_t bytes) { char *q = lds + o; o += (bytes + 15) & ~(size_t)15; return q; };
    M *A4 = (M *)take(sizeof(M) * ld * nb);
    M *A3 = (M *)take(sizeof(M) * nt * nb);
    M *A2 = (M *)take(sizeof(M) * nt * nb);
    M *A1 = (M *)take(sizeof(M) * nt * nt);
    double *fac = (double *)take(sizeof(double) * N);
    int *p4 = (int *)take(sizeof(int) * nb);
    cd *tmp = (cd *)take(sizeof(cd) * nb);
    __shared__ int s_info;
    __shared__ int s_p1[4];
    for (int i = lane; i < ld * nb; i += 64) A4[i] = mzero<M>();
    for (int i = lane; i < nt * nb; i += 64) { A3[i] = mzero<M>(); A2[i] = mzero<M>(); }
    for (int i = lane; i < nt * nt; i += 64) A1[i] = mzero<M>();
    __syncthreads();
    // banded rows + A3 + precond of the banded rows
    for (int n_r = 1 + lane; n_r <= nb; n_r += 64) {
        const int idx = n_r + nt - 1;
        double mx = 0.0;
        for (int nbd = 1; nbd <= W; nbd++) {
            double re = S0[(size_t)idx * W + nbd - 1] - c1 * S1[(size_t)idx * W + nbd - 1];
            double im = S2 ? c2 * S2[(size_t)idx * W + nbd - 1] : 0.0;
            M v = mk_elem<M>(re, im);
            double a = S2 ? hypot(re, im) : fabs(re);
            if (a > mx) mx = a;
            int c = n_r + ku + 1 - nbd;
            if (c <= nb && c >= 1) A4[(kl + nbd - 1) + (size_t)(c - 1) * ld] = mscale(v, p.rnorm);
            if (nbd > ku + 1 && nbd <= ku + 1 + kl) {
                int band = nbd - ku - 1; // 1..kl
                if (n_r <= band && n_r + nt - band >= 1)
                    A3[(n_r - 1) + (size_t)(n_r + nt - band - 1) * nb] = mscale(v, p.rnorm);
            }
        }
        fac[idx] = 1.0 / p.rnorm / mx;
    }
    // tau lines
    for (int n = lane; n < N; n += 64)
        for (int i = 0; i < nt; i++) {
            M v = mk_elem<M>(p.rnorm * tauv[i * N + n], 0.0);
            if (n < nt) A1[i + n * nt] = v;
            else A2[i + (size_t)(n - nt) * nt] = v;
        }
    __syncthreads();
    if (lane < nt) A1[lane] = mscale(A1[lane], p.boundary_fac);
    __syncthreads();
    for (int i = 0; i < nt; i++) {
        double mx = 0.0;
        for (int c = lane; c < nb; c += 64) {
            M a = A2[i + (size_t)c * nt];
            double v;
            if constexpr (sizeof(M) == 16) v = hypot(((cd *)&a)->x, ((cd *)&a)->y);
            else v = fabs(*(double *)&a);
            mx = fmax(mx, v);
        }
        for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
        if (lane == 0) fac[i] = 1.0 / mx;
    }
    __syncthreads();
    for (int t = lane; t < nt * (nt + nb); t += 64) {
        int i = t % nt, c = t / nt;
        if (c < nt) A1[i + c * nt] = mscale(A1[i + c * nt], fac[i]);
        else A2[i + (size_t)(c - nt) * nt] = mscale(A2[i + (size_t)(c - nt) * nt], fac[i]);
    }
    for (int t = lane; t < nt * nb; t += 64) {
        int r = t % nb, c = t / nb;
        A3[r + (size_t)c * nb] = mscale(A3[r + (size_t)c * nb], fac[r + nt]);
    }
    for (int t = lane; t < nb * ld; t += 64) {
        int c = t / ld, rowb = t % ld; // storage row, column c
        int i = rowb - (kl + ku) + c;  // matrix row of this entry
        if (rowb >= kl && i >= 0 && i < nb) A4[rowb + (size_t)c * ld] = mscale(A4[rowb + (size_t)c * ld], fac[i + nt]);
    }
    __syncthreads();
    int info = 0;
    wave_gbtrf<M>(nb, kl, ku, A4, p4, &s_info);
    __syncthreads();
    info = s_info;
    // v = A4^{-1} A3, column by column through a complex LDS scratch vector
    if (info == 0) {
        for (int c = 0; c < nt; c++) {
            for (int i = lane; i < nb; i += 64) {
                M a = A3[i + (size_t)c * nb];
                if constexpr (sizeof(M) == 16) tmp[i] = *(cd *)&a;
                else tmp[i] = mk(*(double *)&a, 0.0);
            }
            __syncthreads();
            wave_gbtrs<M>(nb, kl, ku, A4, ld, p4, tmp);
            __syncthreads();
            for (int i = lane; i < nb; i += 64) {
                if constexpr (sizeof(M) == 16) *(cd *)&A3[i + (size_t)c * nb] = tmp[i];
                else *(double *)&A3[i + (size_t)c * nb] = tmp[i].x;
            }
            __syncthreads();
        }
        // Schur complement A1 <- A1 - A2 A3 (dgemm loop order j, l, i)
        if (lane == 0) {
            for (int jj = 0; jj < nt; jj++)
                for (int l = 0; l < nb; l++) {
                    M temp = mneg(A3[l + (size_t)jj * nb]);
                    if (mnz(temp))
                        for (int i = 0; i < nt; i++)
                            A1[i + jj * nt] = madd(A1[i + jj * nt], mmm(temp, A2[i + (size_t)l * nt]));
                }
            // dgetrf of the nt x nt Schur complement
            for (int jj = 0; jj < nt; jj++) {
                int pv = jj;
                for (int i = jj + 1; i < nt; i++)
                    if (mabs1(A1[i + jj * nt]) > mabs1(A1[pv + jj * nt])) pv = i;
                s_p1[jj] = pv;
                if (mnz(A1[pv + jj * nt])) {
                    if (pv != jj)
                        for (int c = 0; c < nt; c++) { M t = A1[jj + c * nt]; A1[jj + c * nt] = A1[pv + c * nt]; A1[pv + c * nt] = t; }
                    M rp = mrecip(A1[jj + jj * nt]);
                    for (int i = jj + 1; i < nt; i++) A1[i + jj * nt] = mmm(A1[i + jj * nt], rp);
                } else if (info == 0) {
                    info = 1000 + jj + 1;
                }
                for (int c = jj + 1; c < nt; c++) {
                    M t = mneg(A1[jj + c * nt]);
                    for (int i = jj + 1; i < nt; i++) A1[i + c * nt] = madd(A1[i + c * nt], mmm(A1[i + jj * nt], t));
                }
            }
            s_info = info;
        }
        __syncthreads();
        info = s_info;
    }
    // write back
    for (int i = lane; i < ld * nb; i += 64) A4g[i] = A4[i];
    for (int i = lane; i < nt * nb; i += 64) { A3g[i] = A3[i]; A2g[i] = A2[i]; }
    for (int i = lane; i < nt * nt; i += 64) A1g[i] = A1[i];
    for (int i = lane; i < N; i += 64) facg[i] = fac[i];
    for (int i = lane; i < nb; i += 64) p4g[i] = p4[i];
    for (int i = lane; i < nb; i += 64) dinvg[i] = mrecip(A4[(kl + ku) + (size_t)i * ld]);
    if (lane < nt) p1g[lane] = s_p1[lane];
    if (lane == 0) *infog = info;
}

// grid: (3, nm) blocks: field 0 = T, 1 = xi, 2 = psi; 64 threads
__device__ __forceinline__ void build_lhs_body(const PzDev &p, int fld, int mi, int force)
{
    extern __shared__ char lds[];
    if (!force && p.ss->lMat == 0) return;
    const int N = p.N, m = p.idx2m[mi];
    const double wimp = p.ss->wimp_lin[0];
    if (fld < 2) {
        if (fld == 0 && !p.l_heat) return;
        if (fld == 1 && !p.l_chem) return;
        const int nt = PZ_T_NT, kl = PZ_T_KL, ku = PZ_T_KU, ld = PZ_T_LD, nb = N - nt, K = 4;
        const double diff = fld == 0 ? p.TdiffFac * p.hdifT[mi] : p.XidiffFac * p.hdifXi[mi];
        const double c1 = wimp * (fld == 0 ? p.TdiffFac : p.XidiffFac) * (fld == 0 ? p.hdifT[mi] : p.hdifXi[mi]);
        (void)diff;
        const int ktop = fld == 0 ? p.ktopt : p.ktopxi, kbot = fld == 0 ? p.kbott : p.kbotxi;
        double *tauv = (double *)lds;
        for (int n = threadIdx.x; n < N; n += 64) {
            tauv[n] = ktop == 1 ? p.tau[n] : p.tau[2 * N + n];
            tauv[N + n] = kbot == 1 ? p.tau[N + n] : p.tau[3 * N + n];
        }
        __syncthreads();
        const double *S0 = p.T_S0 + (size_t)mi * N * (2 * K + 1);
        const double *S1 = (fld == 0 ? p.T_S1 : p.X_S1) + (size_t)mi * N * (2 * K + 1);
        double *A1 = (fld == 0 ? p.T_A1 : p.X_A1) + (size_t)mi * nt * nt;
        double *A2 = (fld == 0 ? p.T_A2 : p.X_A2) + (size_t)mi * nt * nb;
        double *A3 = (fld == 0 ? p.T_A3 : p.X_A3) + (size_t)mi * nt * nb;
        double *A4 = (fld == 0 ? p.T_A4 : p.X_A4) + (size_t)mi * ld * nb;
        int *p1 = (fld == 0 ? p.T_p1 : p.X_p1) + (size_t)mi * nt;
        int *p4 = (fld == 0 ? p.T_p4 : p.X_p4) + (size_t)mi * nb;
        double *fac = (fld == 0 ? p.T_fac : p.X_fac) + (size_t)mi * N;
        double *dinv = (fld == 0 ? p.T_dinv : p.X_dinv) + (size_t)mi * nb;
        build_bordered<double>(p, nt, kl, ku, K, S0, S1, nullptr, c1, 0.0, tauv, A1, A2, A3, A4, p1, p4, dinv, fac,
                               p.info + fld * p.nm + mi, lds + 4 * N * sizeof(double));
    } else {
        const int nt = m == 0 ? 2 : PZ_P_NT, kl = m == 0 ? 4 : PZ_P_KL, ku = kl;
        const int nb = N - nt, K = PZ_P_KL;
        const double nu = p.ViscFac * p.hdifV[mi];
        const double c1 = wimp * p.ViscFac * p.hdifV[mi];
        const double c2 = (m == 0 || !p.l_coriolis_imp) ? 0.0 : wimp * p.CorFac * (double)m;
        (void)nu;
        double *tauv = (double *)lds;
        const double rc_or1 = p.or1[0], rb_or1 = p.or1[N - 1];
        for (int n = threadIdx.x; n < N; n += 64) {
            double rT = p.tau[n], rB = p.tau[N + n], dT = p.tau[2 * N + n], dB = p.tau[3 * N + n];
            double d2T = p.tau[4 * N + n], d2B = p.tau[5 * N + n];
            double dn2 = (double)n * (double)n;
            double d3top = 8.0 / 15.0 * dn2 * (dn2 - 1.0) * (dn2 - 4.0);
            if (m == 0) {
                tauv[n] = p.ktopv == 1 ? (dT - rc_or1 * rT) : rT;
                tauv[N + n] = p.kbotv == 1 ? (dB - rb_or1 * rB) : rB;
            } else {
                tauv[n] = rT;
                tauv[N + n] = rB;
                tauv[2 * N + n] = p.ktopv == 1 ? (d2T - rc_or1 * dT) : d3top;
                tauv[3 * N + n] = p.kbotv == 1 ? (d2B - rb_or1 * dB) : dB;
            }
        }
        __syncthreads();
        const size_t W = 2 * K + 1;
        const int ldm = 2 * kl + ku + 1;
        // m = 0 stencils are stored 9 wide (K = 4) at the start of the m block
        const double *S0 = p.P_S0 + (size_t)mi * N * W;
        const double *S1 = p.P_S1 + (size_t)mi * N * W;
        const double *S2 = p.P_S2 + (size_t)mi * N * W;
        cd *A1 = p.P_A1 + (size_t)mi * 16;
        cd *A2 = p.P_A2 + (size_t)mi * PZ_P_NT * N;
        cd *A3 = p.P_A3 + (size_t)mi * PZ_P_NT * N;
        cd *A4 = p.P_A4 + (size_t)mi * PZ_P_LD * N;
        int *p1 = p.P_p1 + (size_t)mi * 4;
        int *p4 = p.P_p4 + (size_t)mi * N;
        double *fac = p.P_fac + (size_t)mi * N;
        cd *dinv = p.P_dinv + (size_t)mi * N;
        (void)ldm;
        char *scr = lds + 4 * N * sizeof(double);
        if (m == 0)
            build_bordered<cd>(p, nt, kl, ku, 4, S0, S1, nullptr, c1, 0.0, tauv, A1, A2, A3, A4, p1, p4, dinv,
                               fac, p.info + 2 * p.nm + mi, scr);
        else
            build_bordered<cd>(p, nt, kl, ku, K, S0, S1, S2, c1, c2, tauv, A1, A2, A3, A4, p1, p4, dinv, fac,
                               p.info + 2 * p.nm + mi, scr);
    }
}

__global__ void __launch_bounds__(64) k_build_lhs(const PzDev *__restrict__ pd, int force)
{
    build_lhs_body(*pd, blockIdx.x, blockIdx.y, force);
}
// 1 when any run of a batch rebuilds its matrices in this step (the lanes of
// the first wave read the runs' flags); one dependent load chain instead of
// one per block of the batch
__device__ int batch_any_lmat(const PzBatch &b, int force)
{
    __shared__ int s_any;
    if (threadIdx.x < 64) {
        int any = force;
        for (int r = threadIdx.x; r < b.nruns; r += 64) any |= b.runs[r]->ss->lMat != 0;
        any = __any(any);
        if (threadIdx.x == 0) s_any = any;
    }
    __syncthreads();
    return s_any;
}

// batched: local block index = fld + 3 mi (grid-stride over the nb blocks)
__global__ void __launch_bounds__(64) k_build_lhs_batch(PzBatch b, int nb, int force)
{
    if (!batch_any_lmat(b, force)) return;
    for (int i = blockIdx.x; i < nb; i += gridDim.x) {
        const int2 e = b.map[i];
        build_lhs_body(*b.runs[e.x], e.y % 3, e.y / 3, force);
        __syncthreads();
    }
}

// global -> LDS copy by the whole block, four loads in flight per thread
// global -> LDS copy of a byte range (dst in LDS, src in global memory)
__device__ __forceinline__ void lds_copy(void *dst, const void *src, size_t bytes)
{
    const size_t st = blockDim.x;
    size_t i = threadIdx.x;
    if ((bytes & 7) == 0 && ((size_t)src & 7) == 0) {
        g_double *s = (g_double *)src;
        lds_double *d = (lds_double *)dst;
        const size_t n = bytes / 8;
        for (; i + 3 * st < n; i += 4 * st) {
            const double a = s[i], b = s[i + st], c = s[i + 2 * st], e = s[i + 3 * st];
            d[i] = a; d[i + st] = b; d[i + 2 * st] = c; d[i + 3 * st] = e;
        }
        for (; i < n; i += st) d[i] = s[i];
    } else {
        g_int *s = (g_int *)src;
        __attribute__((address_space(3))) int *d = (__attribute__((address_space(3))) int *)dst;
        const size_t n = bytes / 4;
        for (; i + 3 * st < n; i += 4 * st) {
            const int a = s[i], b = s[i + st], c = s[i + 2 * st], e = s[i + 3 * st];
            d[i] = a; d[i + st] = b; d[i + 2 * st] = c; d[i + 3 * st] = e;
        }
        for (; i < n; i += st) d[i] = s[i];
    }
}

// ---------------------------------------------------------------------------
// Dense solve by explicit inverse: y_i = sum_j INV[j][i] x_j (i < nrow,
// j < ncol) for NMAT matrices at once (column-major, leading dimension
// ldi, columns 16-byte aligned).  E = double: a thread owns two consecutive rows (one 16-byte
// load per column); E = cd: one row.  Threads split the columns into CG
// groups, each thread keeps 16 column loads in flight, and the partial sums
// are combined in a fixed order (bitwise reproducible).  x and y may alias;
// every thread of the block must call it.  red: LDS scratch of
// matvec_scratch(nrow, ncol, blockDim.x, RPT, NMAT) complex values.
// ---------------------------------------------------------------------------
// complex values of scratch for nmat matrices of nrow x ncol with `threads`
// threads, rpt rows per thread: (CG - 1) partial sums of R row units each
// (at most 2 threads nmat)
__host__ __device__ inline size_t matvec_scratch(int nrow, int ncol, int threads, int rpt, int nmat)
{
    const int R = (nrow + rpt - 1) / rpt;
    const int CG = threads / R < ncol ? threads / R : ncol;
    return CG > 1 ? (size_t)nmat * (size_t)(CG - 1) * (size_t)R * (size_t)rpt : 0;
}
template <class E, int NMAT>
__device__ __forceinline__ void block_matvec_rc(int nrow, int ncol, int ldi, VecSet<const E> inv, VecSet<cd> x, VecSet<cd> y,
                                cd *red)
{
    constexpr int RPT = sizeof(E) == 8 ? 2 : 1, U = 16 / NMAT;
    const int N = ncol;
    const int R = (nrow + RPT - 1) / RPT;
    const int CG = min((int)blockDim.x / R, N);
    const int t = threadIdx.x, cg = t / R, ru = t - cg * R;
    const size_t RS = matvec_scratch(nrow, ncol, blockDim.x, RPT, 1); // scratch per matrix
    cd acc[NMAT][RPT];
#pragma unroll
    for (int a = 0; a < NMAT; a++)
#pragma unroll
        for (int r = 0; r < RPT; r++) acc[a][r] = mk(0.0, 0.0);
    if (cg < CG) {
        const int nj = (N - cg + CG - 1) / CG;
        for (int jb = 0; jb < nj; jb += U) {
            double2 v[NMAT][U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int jj = cg + (jb + u) * CG;
                const bool ok = jb + u < nj;
#pragma unroll
                for (int a = 0; a < NMAT; a++)
                    v[a][u] = ok ? gld2((const double *)(inv[a] + (size_t)jj * ldi + (size_t)ru * RPT)) : mk(0, 0);
            }
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int a = 0; a < NMAT; a++) {
                    const int jj = min(cg + (jb + u) * CG, N - 1);
                    const cd xj = lds_load(x[a], jj);
                    if constexpr (RPT == 2) {
                        acc[a][0].x = fma(v[a][u].x, xj.x, acc[a][0].x);
                        acc[a][0].y = fma(v[a][u].x, xj.y, acc[a][0].y);
                        acc[a][1].x = fma(v[a][u].y, xj.x, acc[a][1].x);
                        acc[a][1].y = fma(v[a][u].y, xj.y, acc[a][1].y);
                    } else {
                        acc[a][0] = cadd(acc[a][0], cmul(v[a][u], xj));
                    }
                }
        }
        if (cg > 0)
#pragma unroll
            for (int a = 0; a < NMAT; a++)
#pragma unroll
                for (int r = 0; r < RPT; r++) lds_store(red, (size_t)a * RS + (cg - 1) * R * RPT + ru * RPT + r, acc[a][r]);
    }
    __syncthreads();
    if (cg == 0) {
#pragma unroll
        for (int a = 0; a < NMAT; a++)
#pragma unroll
            for (int r = 0; r < RPT; r++) {
                cd s = acc[a][r];
                for (int c = 1; c < CG; c++) s = cadd(s, lds_load(red, (size_t)a * RS + (c - 1) * R * RPT + ru * RPT + r));
                const int row = ru * RPT + r;
                if (row < nrow) lds_store(y[a], row, s);
            }
    }
    __syncthreads();
}
// square n x n form
template <class E, int NMAT>
__device__ __forceinline__ void block_matvec(int N, int ldi, VecSet<const E> inv, VecSet<cd> x, VecSet<cd> y,
                                             cd *red)
{
    block_matvec_rc<E, NMAT>(N, N, ldi, inv, x, y, red);
}

// LDS layout of k_build_inv: one field's bordered factors (sized for the
// largest, psi with m != 0) and one right-hand side per wave.
struct InvLayout {
    size_t A4, A2, A3, A1, p4, p1, dinv, rhs, total;
    __host__ __device__ InvLayout(int N, int nw)
    {
        size_t o = 0;
        auto take = [&](size_t bytes) { size_t r = o; o += (bytes + 15) & ~(size_t)15; return r; };
        A4 = take(sizeof(cd) * PZ_P_LD * N);
        A2 = take(sizeof(cd) * PZ_P_NT * N);
        A3 = take(sizeof(cd) * PZ_P_NT * N);
        A1 = take(sizeof(cd) * 16);
        p4 = take(sizeof(int) * N);
        p1 = take(sizeof(int) * 4);
        dinv = take(sizeof(cd) * N);
        rhs = take(sizeof(cd) * (size_t)nw * N);
        total = o;
    }
};

// Columns of the inverses INV = A^{-1} diag(fac) of the bordered systems
// factorised by k_build_lhs: wave w solves A y = fac_j e_j for the columns
// j = w, w + nw, ... with the same bordered solve (algebra_lapack.f90:332-416)
// as one time step would.  grid (nm, 3): field 0 = T, 1 = xi, 2 = psi.
__device__ __forceinline__ void build_inv_body(const PzDev &p, int mi, int fld, int force)
{
    extern __shared__ char lds[];
    if (!force && p.ss->lMat == 0) return;
    const int N = p.N, m = p.idx2m[mi];
    if ((fld == 0 && !p.l_heat) || (fld == 1 && !p.l_chem)) return;
    const int nw = blockDim.x >> 6, wave = threadIdx.x >> 6, lane = threadIdx.x & 63, ldi = p.ldi;
    const InvLayout lay(N, nw);
    cd *rhs = (cd *)(lds + lay.rhs) + (size_t)wave * N;
    if (fld < 2) {
        const int nt = PZ_T_NT, nb = N - nt;
        const double *fac = (fld == 0 ? p.T_fac : p.X_fac) + (size_t)mi * N;
        lds_copy(lds + lay.A4, (fld == 0 ? p.T_A4 : p.X_A4) + (size_t)mi * PZ_T_LD * nb, sizeof(double) * PZ_T_LD * nb);
        lds_copy(lds + lay.A2, (fld == 0 ? p.T_A2 : p.X_A2) + (size_t)mi * nt * nb, sizeof(double) * nt * nb);
        lds_copy(lds + lay.A3, (fld == 0 ? p.T_A3 : p.X_A3) + (size_t)mi * nt * nb, sizeof(double) * nt * nb);
        lds_copy(lds + lay.A1, (fld == 0 ? p.T_A1 : p.X_A1) + (size_t)mi * nt * nt, sizeof(double) * nt * nt);
        lds_copy(lds + lay.p4, (fld == 0 ? p.T_p4 : p.X_p4) + (size_t)mi * nb, sizeof(int) * nb);
        lds_copy(lds + lay.p1, (fld == 0 ? p.T_p1 : p.X_p1) + (size_t)mi * nt, sizeof(int) * nt);
        lds_copy(lds + lay.dinv, (fld == 0 ? p.T_dinv : p.X_dinv) + (size_t)mi * nb, sizeof(double) * nb);
        __syncthreads();
        double *inv = (fld == 0 ? p.T_inv : p.X_inv) + (size_t)mi * N * ldi;
        for (int j = wave; j < N; j += nw) {
            for (int k = lane; k < N; k += 64) rhs[k] = mk(k == j ? fac[j] : 0.0, 0.0);
            __builtin_amdgcn_wave_barrier();
            wave_bordered_fast<double, PZ_T_KL, PZ_T_KU>(N, nt, (double *)(lds + lay.A1), (int *)(lds + lay.p1),
                                                         (double *)(lds + lay.A2), (double *)(lds + lay.A3),
                                                         (double *)(lds + lay.A4), (int *)(lds + lay.p4),
                                                         (double *)(lds + lay.dinv), rhs);
            for (int k = lane; k < ldi; k += 64) inv[(size_t)j * ldi + k] = k < N ? rhs[k].x : 0.0;
            __builtin_amdgcn_wave_barrier();
        }
    } else {
        const int nt = m == 0 ? 2 : PZ_P_NT, kl = m == 0 ? 4 : PZ_P_KL, ld = 3 * kl + 1, nb = N - nt;
        const double *fac = p.P_fac + (size_t)mi * N;
        lds_copy(lds + lay.A4, p.P_A4 + (size_t)mi * PZ_P_LD * N, sizeof(cd) * ld * nb);
        lds_copy(lds + lay.A2, p.P_A2 + (size_t)mi * PZ_P_NT * N, sizeof(cd) * nt * nb);
        lds_copy(lds + lay.A3, p.P_A3 + (size_t)mi * PZ_P_NT * N, sizeof(cd) * nt * nb);
        lds_copy(lds + lay.A1, p.P_A1 + (size_t)mi * 16, sizeof(cd) * nt * nt);
        lds_copy(lds + lay.p4, p.P_p4 + (size_t)mi * N, sizeof(int) * nb);
        lds_copy(lds + lay.p1, p.P_p1 + (size_t)mi * 4, sizeof(int) * nt);
        lds_copy(lds + lay.dinv, p.P_dinv + (size_t)mi * N, sizeof(cd) * nb);
        __syncthreads();
        cd *A1 = (cd *)(lds + lay.A1), *A2 = (cd *)(lds + lay.A2), *A3 = (cd *)(lds + lay.A3);
        cd *A4 = (cd *)(lds + lay.A4), *di = (cd *)(lds + lay.dinv);
        int *p1 = (int *)(lds + lay.p1), *p4 = (int *)(lds + lay.p4);
        cd *inv = p.P_inv + (size_t)mi * N * ldi;
        for (int j = wave; j < N; j += nw) {
            for (int k = lane; k < N; k += 64) rhs[k] = mk(k == j ? fac[j] : 0.0, 0.0);
            __builtin_amdgcn_wave_barrier();
            if (m == 0) wave_bordered_fast<cd, 4, 4>(N, nt, A1, p1, A2, A3, A4, p4, di, rhs);
            else wave_bordered_fast<cd, PZ_P_KL, PZ_P_KU>(N, nt, A1, p1, A2, A3, A4, p4, di, rhs);
            for (int k = lane; k < ldi; k += 64) inv[(size_t)j * ldi + k] = k < N ? rhs[k] : mk(0.0, 0.0);
            __builtin_amdgcn_wave_barrier();
        }
    }
}

__global__ void __launch_bounds__(1024) k_build_inv(const PzDev *__restrict__ pd, int force)
{
    build_inv_body(*pd, blockIdx.x, blockIdx.y, force);
}
// batched: local block index = mi + nm fld (grid-stride over the nb blocks)
__global__ void __launch_bounds__(1024) k_build_inv_batch(PzBatch b, int nb, int force)
{
    if (!batch_any_lmat(b, force)) return;
    for (int i = blockIdx.x; i < nb; i += gridDim.x) {
        const int2 e = b.map[i];
        const PzDev &p = *b.runs[e.x];
        build_inv_body(p, e.y % p.nm, e.y / p.nm, force);
        __syncthreads();
    }
}
