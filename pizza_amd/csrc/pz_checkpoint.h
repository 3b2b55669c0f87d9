// Host side of libpizza_amd: checkpoint and frame files.
//
// Checkpoints follow checkpoints.f90: version 7 is written
// (write_checkpoint_mloc, checkpoints.f90:33-183) and versions 1-7 are read
// (read_checkpoint, :185-707).  A file is a header followed by complex(8)
// arrays f(n_m_max, n_r_max) stored m fastest (Fortran order): u_s, u_phi
// and the psi history, then T and its history, then xi and its history
// (then the phase field, which this path ignores).  The history of a
// multistep scheme is expl(2:nexp), impl(2:nimp), old(2:nold) in the
// representation of the run's radial method (grid values for collocation
// and finite differences, Chebyshev space for the integration method).
// Frames follow write_snapshot_mloc (output_frames.f90:96-167).
#pragma once
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace pzck {

typedef double2 c2;

// Sequential reader of a stream-access file; throws on a short read.
class Reader {
  public:
    explicit Reader(const char *path) : f_(std::fopen(path, "rb"))
    {
        if (!f_) throw std::runtime_error(std::string("cannot open ") + path);
    }
    ~Reader()
    {
        if (f_) std::fclose(f_);
    }
    template <class T> T get()
    {
        T v;
        raw(&v, sizeof(T));
        return v;
    }
    void raw(void *p, size_t n)
    {
        if (std::fread(p, 1, n, f_) != n) throw std::runtime_error("checkpoint file is truncated");
        pos_ += n;
    }
    std::string text(size_t n)
    {
        std::string s(n, ' ');
        raw(&s[0], n);
        return s;
    }
    long long pos() const { return pos_; }

  private:
    FILE *f_;
    long long pos_ = 0;
};

class Writer {
  public:
    explicit Writer(const char *path) : f_(std::fopen(path, "wb"))
    {
        if (!f_) throw std::runtime_error(std::string("cannot create ") + path);
    }
    ~Writer()
    {
        if (f_) std::fclose(f_);
    }
    template <class T> void put(const T &v) { raw(&v, sizeof(T)); }
    void raw(const void *p, size_t n)
    {
        if (std::fwrite(p, 1, n, f_) != n) throw std::runtime_error("short write of a checkpoint file");
    }
    void text(const std::string &s, size_t n)
    {
        std::string t = s;
        t.resize(n, ' ');
        raw(t.data(), n);
    }
    void close()
    {
        if (f_ && std::fclose(f_) != 0) {
            f_ = nullptr;
            throw std::runtime_error("closing a checkpoint file failed");
        }
        f_ = nullptr;
    }

  private:
    FILE *f_;
};

// Everything the header of a checkpoint says (all versions).
struct Header {
    int version = 0;
    double time = 0.0;
    bool l_coll = true;
    std::string family = "MULTISTEP";
    int nexp = 2, nimp = 2, nold = 2; // as stored (before the version <= 5 shift)
    std::vector<double> dt;           // as stored
    double ra = 0, pr = 0, raxi = 0, sc = 0, ek = 0, stef = 0, radratio = 0;
    int n_r = 0, m_max = 0, minc = 1;
    std::string rscheme = "cheb";
    int n_max = 0, order_boundary = 0;
    double alph1 = 0, alph2 = 0;
    std::vector<double> r;
    bool l_heat = true, l_chem = false, l_phase = false;
    long long header_bytes = 0;

    bool dirk() const { return family.compare(0, 4, "DIRK") == 0; }
    bool cheb() const { return rscheme.compare(0, 4, "cheb") == 0; }
    int n_m() const { return m_max / minc + 1; }
    // history arrays that follow each field (read_checkpoint's loops)
    int nhist() const
    {
        if (dirk()) return 0;
        int ne = nexp, ni = nimp, no = nold;
        if (version <= 5) { ni--; no--; }
        return (ne - 1) + (ni - 1) + (version > 3 ? no - 1 : 0);
    }
    int n_arrays() const { return (2 + nhist()) + (l_heat ? 1 + nhist() : 0) + (l_chem ? 1 + nhist() : 0) +
                                  (l_phase ? 1 + nhist() : 0); }
};

// Fortran logical as stored by the writing compiler: any nonzero is true.
inline bool logical(int v) { return v != 0; }

// Reads the header, leaving the reader at the first field.
inline Header read_header(Reader &in)
{
    Header h;
    h.version = in.get<int>();
    if (h.version == 1) {
        // CN/AB2 files: time, dt(2), dt(1)
        h.time = in.get<double>();
        double d2 = in.get<double>(), d1 = in.get<double>();
        h.dt = {d1, d2};
        h.nexp = h.nimp = h.nold = 2;
    } else if (h.version == 2 || h.version == 3) {
        h.time = in.get<double>();
        if (h.version == 3) h.l_coll = logical(in.get<int>());
        h.nexp = in.get<int>();
        h.nold = in.get<int>();
        h.nimp = h.nold;
    } else if (h.version == 4) {
        h.time = in.get<double>();
        h.l_coll = logical(in.get<int>());
        h.nexp = in.get<int>();
        h.nimp = in.get<int>();
        h.nold = in.get<int>();
    } else if (h.version >= 5 && h.version <= 7) {
        h.time = in.get<double>();
        h.l_coll = logical(in.get<int>());
        h.family = in.text(10);
        h.nexp = in.get<int>();
        h.nimp = in.get<int>();
        h.nold = in.get<int>();
    } else {
        throw std::runtime_error("unknown checkpoint version " + std::to_string(h.version));
    }
    if (h.nexp < 1 || h.nexp > 16 || h.nimp < 0 || h.nimp > 16 || h.nold < 0 || h.nold > 16)
        throw std::runtime_error("implausible time-scheme sizes in the checkpoint header");
    if (h.version >= 2) {
        h.dt.resize(h.dirk() ? 1 : h.nexp);
        for (double &v : h.dt) v = in.get<double>();
    }
    h.ra = in.get<double>();
    h.pr = in.get<double>();
    h.raxi = in.get<double>();
    h.sc = in.get<double>();
    h.ek = in.get<double>();
    if (h.version >= 7) h.stef = in.get<double>();
    h.radratio = in.get<double>();
    h.n_r = in.get<int>();
    h.m_max = in.get<int>();
    h.minc = in.get<int>();
    if (h.n_r < 4 || h.n_r > 1000000 || h.m_max < 0 || h.m_max > 1000000 || h.minc < 1)
        throw std::runtime_error("implausible grid sizes in the checkpoint header");
    h.rscheme = in.text(72);
    h.n_max = in.get<int>();
    h.order_boundary = in.get<int>();
    h.alph1 = in.get<double>();
    h.alph2 = in.get<double>();
    h.r.resize(h.n_r);
    in.raw(h.r.data(), sizeof(double) * h.r.size());
    h.l_heat = logical(in.get<int>());
    h.l_chem = logical(in.get<int>());
    if (h.version >= 7) h.l_phase = logical(in.get<int>());
    h.header_bytes = in.pos();
    return h;
}

// costf1 of a real vector of length n: the DCT-I with the sqrt(0.5/(n-1))
// scaling of dct_fftw.f90:231-246, as its direct cosine sum in long double
inline void costf(std::vector<double> &f, int n)
{
    const long double pi = 3.14159265358979323846264338327950288L;
    std::vector<double> g(n);
    for (int k = 0; k < n; k++) {
        long double acc = (long double)f[0] + ((k & 1) ? -(long double)f[n - 1] : (long double)f[n - 1]);
        for (int j = 1; j < n - 1; j++)
            acc += 2.0L * f[j] * cosl(pi * (long double)(((long)j * k) % (2L * (n - 1))) / (long double)(n - 1));
        g[k] = (double)(acc * sqrtl(0.5L / (long double)(n - 1)));
    }
    for (int k = 0; k < n; k++) f[k] = g[k];
}

// The grid a field is mapped onto.
struct Target {
    int N = 0, nm = 0;
    const int *idx2m = nullptr;
    const double *r = nullptr;
    bool cheb = true;
    int order_boundary = 0;
    double rnorm = 1.0; // sqrt(2/(N-1)) on Chebyshev grids
};

// map_field / map_field_r (checkpoints.f90:709-866): one stored array
// (n_m_old, n_r_old), m fastest, onto the target in M-layout [m][r].
// Same radial grid: a copy per m; both Chebyshev with the same mapping:
// through the coefficients (costf, zero padding or truncation, costf,
// rescaling); otherwise the cubic through the four stored radii around each
// new radius.  lBc with grid values first replaces the outer boundary value
// by the linear extrapolation of the next two (beta is singular there).
inline void map_field(const Header &hd, const std::vector<c2> &old, const Target &t, bool lBc, bool l_phys,
                      std::vector<c2> &out)
{
    const int No = hd.n_r, nmo = hd.n_m(), N = t.N;
    out.assign((size_t)t.nm * N, c2{0.0, 0.0});
    const bool same = No == N && hd.order_boundary == t.order_boundary && hd.cheb() == t.cheb;
    const bool spectral = hd.cheb() && t.cheb && hd.order_boundary == t.order_boundary;
    std::vector<c2> col(No);
    std::vector<double> re(std::max(N, No)), im(std::max(N, No));
    for (int i = 0; i < t.nm; i++) {
        const int m = t.idx2m[i];
        if (m > hd.m_max || m % hd.minc) continue; // new modes start from zero
        const int io = m / hd.minc;
        for (int k = 0; k < No; k++) col[k] = old[(size_t)k * nmo + io];
        c2 *d = out.data() + (size_t)i * N;
        if (same) {
            for (int k = 0; k < N; k++) d[k] = col[k];
            continue;
        }
        if (lBc && l_phys) col[0] = c2{2.0 * col[1].x - col[2].x, 2.0 * col[1].y - col[2].y};
        if (spectral) {
            for (int k = 0; k < No; k++) { re[k] = col[k].x; im[k] = col[k].y; }
            if (l_phys) { costf(re, No); costf(im, No); }
            for (int k = No; k < N; k++) re[k] = im[k] = 0.0;
            if (l_phys) { costf(re, N); costf(im, N); }
            const double fac = std::sqrt(2.0 / (double)(No - 1)) / t.rnorm;
            for (int k = 0; k < N; k++) d[k] = c2{fac * re[k], fac * im[k]};
        } else {
            for (int k = 0; k < N; k++) {
                const double x = t.r[k];
                int j = 0;
                for (int q = 1; q < No; q++)
                    if (std::fabs(hd.r[q] - x) < std::fabs(hd.r[j] - x)) j = q;
                j = std::min(std::max(j, 2), No - 2);
                // Lagrange form of the cubic through radii j-2 .. j+1
                double sx = 0.0, sy = 0.0;
                for (int a = j - 2; a <= j + 1; a++) {
                    double w = 1.0;
                    for (int b = j - 2; b <= j + 1; b++)
                        if (b != a) w *= (x - hd.r[b]) / (hd.r[a] - hd.r[b]);
                    sx += w * col[a].x;
                    sy += w * col[a].y;
                }
                d[k] = c2{sx, sy};
            }
        }
    }
}

} // namespace pzck
