// libpizza_amd: host side of the MI355X-native IMEX step and its C ABI
// (include/pizza_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pizza_amd.h"
#include "pz_host_ops.h"
#include "pz_mphase.h"
#include "pz_coll.h"
#include "pz_fd.h"
#include "pz_assembly.h"
#include "pz_adjoint.h"
#include "pz_checkpoint.h"

using pzh::ld;

// Staging memory of pz_get_ddr_noalloc: three buffers of PZ_SCRATCH_ELEMS
// complex values (12 MiB) that every GPU running the library holds until the
// process ends.  Host code never passes the array's name to a kernel: it
// takes the device address of this GPU's copy from scratch_address().
#define PZ_SCRATCH_ELEMS (1 << 18)
static __device__ __align__(256) double2 scratch[3 * PZ_SCRATCH_ELEMS];

// The device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy at its own address).
static double2 *scratch_address()
{
    static std::mutex mu;
    static double2 *addr[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) throw std::runtime_error("hipGetDevice failed");
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        addr[dev] = (double2 *)p;
    }
    return addr[dev];
}

// Layout changes between the reference's f(n_m, n_r) (m fastest) and the
// M-layout [m][r] of the transform kernels.
__global__ void __launch_bounds__(256) k_from_ref_layout(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                          int n_m, int N)
{
    const size_t S = (size_t)n_m * N;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < S; q += (size_t)gridDim.x * blockDim.x) {
        const size_t i = q / N, k = q % N;
        dst[q] = src[k * n_m + i];
    }
}
__global__ void __launch_bounds__(256) k_to_ref_layout(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                        int n_m, int N)
{
    const size_t S = (size_t)n_m * N;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < S; q += (size_t)gridDim.x * blockDim.x) {
        const size_t k = q / n_m, i = q % n_m;
        dst[q] = src[i * N + k];
    }
}

namespace {

struct Scheme {
    int id = PZ_CNAB2, family = 0, nstages = 1, nexp = 2, nimp = 1, nold = 1;
    double courfac = 0, wimp_lin0 = 0;
    double bimp0[PZ_MAX_STAGES][PZ_MAX_STAGES] = {}, bexp0[PZ_MAX_STAGES][PZ_MAX_STAGES] = {};
    int l_exp_calc[8] = {}, l_imp_calc_rhs[8] = {};
    int l_assembly = 0;
    double bass_imp0[PZ_MAX_STAGES] = {}, bass_exp0[PZ_MAX_STAGES] = {};
};

// Butcher tables of a DIRK scheme without an assembly stage: n = nstages + 1
// rows, row-major; the implicit diagonal coefficient is bi[1][1].
void set_butcher(Scheme &s, int nst, const double *bi, const double *be)
{
    const int n = nst + 1;
    s.nstages = s.nimp = s.nexp = nst;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) { s.bimp0[i][j] = bi[i * n + j]; s.bexp0[i][j] = be[i * n + j]; }
    for (int k = 0; k < nst; k++) s.l_exp_calc[k] = s.l_imp_calc_rhs[k] = 1;
}

// Tables of a DIRK scheme with an assembly stage: nst x nst Butcher tables
// (dirk_schemes.f90:239-243) and the assembly weights.
void set_butcher_ass(Scheme &s, int nst, const double *bi, const double *be, const double *bai, const double *bae)
{
    s.nstages = s.nimp = s.nexp = nst;
    for (int i = 0; i < nst; i++) {
        for (int j = 0; j < nst; j++) { s.bimp0[i][j] = bi[i * nst + j]; s.bexp0[i][j] = be[i * nst + j]; }
        s.bass_imp0[i] = bai[i];
        s.bass_exp0[i] = bae[i];
    }
    for (int k = 0; k < nst; k++) s.l_exp_calc[k] = s.l_imp_calc_rhs[k] = 1;
    s.l_assembly = 1;
}

// select_time_scheme / initialize of multistep_schemes.f90:40-135 and
// dirk_schemes.f90:34-263, with the tables of set_weights
// (dirk_schemes.f90:286-525) for the schemes without an assembly stage
bool make_scheme(int id, double courfac_nml, Scheme &s)
{
    double cf;
    s.id = id;
    s.nold = 1;
    if (id >= PZ_CNAB2 && id <= PZ_TVB33) {
        s.family = 0;
        s.nstages = 1;
        s.l_exp_calc[0] = s.l_imp_calc_rhs[0] = 1;
        switch (id) {
        case PZ_CNAB2: s.nimp = 1; s.nold = 1; s.nexp = 2; cf = 3.0; break;
        case PZ_MODCNAB: s.nimp = 2; s.nold = 2; s.nexp = 2; cf = 2.9; break;
        case PZ_CNLF: s.nimp = 2; s.nold = 2; s.nexp = 2; cf = 3.0; break;
        case PZ_SBDF2: s.nimp = 1; s.nold = 2; s.nexp = 2; s.l_imp_calc_rhs[0] = 0; cf = 2.8; break;
        case PZ_SBDF3: s.nimp = 1; s.nold = 3; s.nexp = 3; s.l_imp_calc_rhs[0] = 0; cf = 4.0; break;
        case PZ_TVB33: s.nimp = 3; s.nold = 3; s.nexp = 3; cf = 4.0; break;
        default: s.nimp = 1; s.nold = 4; s.nexp = 4; s.l_imp_calc_rhs[0] = 0; cf = 5.5; break; // SBDF4
        }
    } else {
        s.family = 1;
        const double r2 = std::sqrt(2.0);
        switch (id) {
        case PZ_ARS222: {
            const double g = 0.5 * (2.0 - r2), del = 1.0 - 0.5 / g;
            const double bi[9] = {0, 0, 0, 0, g, 0, 0, 1 - g, g}, be[9] = {0, 0, 0, g, 0, 0, del, 1 - del, 0};
            set_butcher(s, 2, bi, be);
            s.wimp_lin0 = g;
            s.l_imp_calc_rhs[0] = 0;
            cf = 1.35;
            break;
        }
        case PZ_ARS443: {
            const double bi[25] = {0, 0, 0, 0, 0, 0, 0.5, 0, 0, 0, 0, 1.0 / 6.0, 0.5, 0, 0,
                                   0, -0.5, 0.5, 0.5, 0, 0, 1.5, -1.5, 0.5, 0.5};
            const double be[25] = {0, 0, 0, 0, 0, 0.5, 0, 0, 0, 0, 11.0 / 18.0, 1.0 / 18.0, 0, 0, 0,
                                   5.0 / 6.0, -5.0 / 6.0, 0.5, 0, 0, 0.25, 1.75, 0.75, -1.75, 0};
            set_butcher(s, 4, bi, be);
            s.wimp_lin0 = 0.5;
            s.l_imp_calc_rhs[0] = 0;
            cf = 0.9;
            break;
        }
        case PZ_BPR353: {
            const double bi[25] = {0, 0, 0, 0, 0, 0.5, 0.5, 0, 0, 0, 5.0 / 18.0, -1.0 / 9.0, 0.5, 0, 0,
                                   0.5, 0, 0, 0.5, 0, 0.25, 0, 0.75, -0.5, 0.5};
            const double be[25] = {0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 4.0 / 9.0, 2.0 / 9.0, 0, 0, 0,
                                   0.25, 0, 0.75, 0, 0, 0.25, 0, 0.75, 0, 0};
            set_butcher(s, 4, bi, be);
            s.wimp_lin0 = 0.5;
            s.l_exp_calc[3] = 0;
            cf = 1.0;
            break;
        }
        case PZ_DP1_242: {
            const double bi[25] = {0, 0, 0, 0, 0, 0, 0.5, 0, 0, 0, 0, 1.0 / 6.0, 0.5, 0, 0,
                                   0, -0.5, 0.5, 0.5, 0, 0, 1.5, -1.5, 0.5, 0.5};
            const double be[25] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1.0 / 3.0, 0, 0, 0,
                                   0, 1, 0, 0, 0, 0, 0.5, 0, 0.5, 0};
            set_butcher(s, 4, bi, be);
            s.wimp_lin0 = 0.5;
            s.l_imp_calc_rhs[0] = 0;
            s.l_exp_calc[0] = s.l_exp_calc[2] = 0;
            cf = 1.35;
            break;
        }
        case PZ_DP2_242: {
            const double g = 1.0 / 3.0;
            const double bi[25] = {0, 0, 0, 0, 0, 0, g, 0, 0, 0, 0, -g, g, 0, 0,
                                   0, 0, 1 - g, g, 0, 0, 0, 0.5, 0.5 - g, g};
            const double be[25] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                                   0, 0, 1, 0, 0, 0, 0, 0.5, 0.5, 0};
            set_butcher(s, 4, bi, be);
            s.wimp_lin0 = g;
            s.l_imp_calc_rhs[0] = 0;
            s.l_exp_calc[0] = 0;
            cf = 1.35;
            break;
        }
        case PZ_LZ232: {
            const double bi[9] = {0, 0, 0, -0.25, 0.5, 0, 0.5, 0, 0.5}, be[9] = {0, 0, 0, 0.25, 0, 0, -1, 2, 0};
            set_butcher(s, 2, bi, be);
            s.wimp_lin0 = 0.5;
            cf = 1.25;
            break;
        }
        case PZ_CK232: {
            const double g = 1.0 - 0.5 * r2;
            const double bi[9] = {0, 0, 0, -1.0 / 3.0 + 0.5 * r2, g, 0, 0.75 - 0.25 * r2, -0.75 + 0.75 * r2, g};
            const double be[9] = {0, 0, 0, 2.0 / 3.0, 0, 0, 0.25, 0.75, 0};
            set_butcher(s, 2, bi, be);
            s.wimp_lin0 = g;
            cf = 1.25;
            break;
        }
        case PZ_LZ453: {
            const double bi[25] = {0, 0, 0, 0, 0, -44.0 / 45.0, 1.2, 0, 0, 0, -47.0 / 300.0, -0.71, 1.2, 0, 0,
                                   3.375, -3.25, -59.0 / 120.0, 1.2, 0,
                                   89.0 / 50.0, -486.0 / 55.0, 8.9, -562.0 / 275.0, 1.2};
            const double be[25] = {0, 0, 0, 0, 0, 2.0 / 9.0, 0, 0, 0, 0, 71.0 / 420.0, 23.0 / 140.0, 0, 0, 0,
                                   -281.0 / 336.0, 187.0 / 112.0, 0, 0, 0, 0.1, 0, 0.5, 0.4, 0};
            set_butcher(s, 4, bi, be);
            s.wimp_lin0 = 1.2;
            cf = 1.15;
            break;
        }
        case PZ_PC2: {
            const double bi[16] = {0, 0, 0, 0, 0.5, 0.5, 0, 0, 0.5, 0, 0.5, 0, 0.5, 0, 0, 0.5};
            const double be[16] = {0, 0, 0, 0, 1, 0, 0, 0, 0.5, 0.5, 0, 0, 0.5, 0, 0.5, 0};
            set_butcher(s, 3, bi, be);
            s.wimp_lin0 = 0.5;
            cf = 0.8;
            break;
        }
        case PZ_SSP222: { /* :299-311 */
            const double g = 1.0 - 1.0 / r2;
            double bi[9] = {0, 0, 0, 0, g, 0, 0, 1 - 2 * g, g}, be[9] = {0, 0, 0, 0, 0, 0, 0, 1, 0};
            double ba[3] = {0, 0.5, 0.5};
            set_butcher_ass(s, 3, bi, be, ba, ba);
            s.wimp_lin0 = g; s.l_imp_calc_rhs[0] = 0; cf = 1.35;
            break;
        }
        case PZ_LSDIRK222: { /* :312-326 */
            const double g = 1.0 - 1.0 / r2;
            double chat = 1.0 / (2.0 * g);
            double bi[9] = {0, 0, 0, 0, g, 0, 0, 1 - g, g}, be[9] = {0, 0, 0, 0, 0, 0, 0, chat, 0};
            double ba[3] = {0, 1 - g, g};
            set_butcher_ass(s, 3, bi, be, ba, ba);
            s.wimp_lin0 = g; s.l_imp_calc_rhs[0] = 0; cf = 1.35;
            break;
        }
        case PZ_ARS232: { /* :427-441 */
            const double g = 0.5 * (2.0 - r2);
            double del = -2.0 * r2 / 3.0;
            double bi[9] = {0, 0, 0, 0, g, 0, 0, 1 - g, g}, be[9] = {0, 0, 0, g, 0, 0, del, 1 - del, 0};
            double ba[3] = {0, 1 - g, g};
            set_butcher_ass(s, 3, bi, be, ba, ba);
            s.wimp_lin0 = g; s.l_imp_calc_rhs[0] = 0; cf = 1.35;
            break;
        }
        case PZ_ARS233: { /* :327-341 */
            const double g = (3.0 + sqrt(3.0)) / 6.0;
            double bi[9] = {0, 0, 0, 0, g, 0, 0, 1 - 2 * g, g}, be[9] = {0, 0, 0, g, 0, 0, g - 1, 2 * (1 - g), 0};
            double ba[3] = {0, 0.5, 0.5};
            set_butcher_ass(s, 3, bi, be, ba, ba);
            s.wimp_lin0 = g; s.l_imp_calc_rhs[0] = 0; cf = 1.35;
            break;
        }
        case PZ_SSP322: { /* :342-356 */
            double bi[16] = {0, 0, 0, 0, 0, 0.5, 0, 0, 0, -0.5, 0.5, 0, 0, 0, 0.5, 0.5};
            double be[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0};
            double ba[4] = {0, 0, 0.5, 0.5};
            set_butcher_ass(s, 4, bi, be, ba, ba);
            s.wimp_lin0 = 0.5; s.l_imp_calc_rhs[0] = 0; cf = 1.35;
            break;
        }
        case PZ_SSP332: { /* :357-372 */
            const double g = 1.0 - 1.0 / r2;
            double bi[16] = {0, 0, 0, 0, 0, g, 0, 0, 0, 1 - 2 * g, g, 0, 0, 0.5 - g, 0, g};
            double be[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0.25, 0.25, 0};
            double ba[4] = {0, 1.0 / 6.0, 1.0 / 6.0, 2.0 * (1.0 / 3.0)};
            set_butcher_ass(s, 4, bi, be, ba, ba);
            s.wimp_lin0 = g; s.l_imp_calc_rhs[0] = 0; cf = 1.35;
            break;
        }
        case PZ_SSP433: { /* :373-392; alpha, beta, eta are default-real
                             (single precision) literals in the reference */
            const double al = (double)0.24169426078821f, be_ = (double)0.06042356519705f;
            const double et = (double)0.12915286960590f;
            double bi[25] = {0, 0, 0, 0, 0, 0, al, 0, 0, 0, 0, -al, al, 0, 0,
                             0, 0, 1 - al, al, 0, 0, be_, et, 0.5 - be_ - et - al, al};
            double be[25] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                             0, 0, 1, 0, 0, 0, 0, 0.25, 0.25, 0};
            double ba[5] = {0, 0, 1.0 / 6.0, 1.0 / 6.0, 2.0 * (1.0 / 3.0)};
            set_butcher_ass(s, 5, bi, be, ba, ba);
            s.wimp_lin0 = al; s.l_imp_calc_rhs[0] = 0; cf = 1.35;
            break;
        }
        case PZ_ARS343: { /* :526-545 */
            const double g = 0.435866521508459;
            const double b1 = -1.5 * g * g + 4.0 * g - 0.25, b2 = 1.5 * g * g - 5.0 * g + 1.25;
            double bi[16] = {0, 0, 0, 0, 0, g, 0, 0, 0, 0.5 * (1 - g), g, 0, 0, b1, b2, g};
            double be[16] = {0, 0, 0, 0, g, 0, 0, 0, 0.3212788860286278, 0.3966543747256017, 0, 0,
                             -0.1058582960718797, 0.5529291480359398, 0.5529291480359398, 0};
            double ba[4] = {0, b1, b2, g};
            set_butcher_ass(s, 4, bi, be, ba, ba);
            s.wimp_lin0 = g; s.l_imp_calc_rhs[0] = 0; cf = 0.7;
            break;
        }
        case PZ_MARS343: { /* :571-590 */
            const double g = 0.435866521508458;
            const double b1 = 1.20849664917601, b2 = -0.64436317068446;
            double bi[16] = {0, 0, 0, 0, 0, g, 0, 0, 0, 0.5 * (1 - g), g, 0, 0, b1, b2, g};
            double be[16] = {0, 0, 0, 0, 0.877173304301691, 0, 0, 0, 0.535396540307354, 0.182536720446875, 0, 0,
                             0.63041255815287, -0.83193390106308, 1.20152134291021, 0};
            double ba[4] = {0, b1, b2, g};
            set_butcher_ass(s, 4, bi, be, ba, ba);
            s.wimp_lin0 = g; s.l_imp_calc_rhs[0] = 0; cf = 0.5;
            break;
        }
        case PZ_KC343: { /* :546-570 */
            const double g = 1767732205903.0 / 4055673282236.0;
            const double a30 = 1471266399579.0 / 7840856788654.0, a31 = -4482444167858.0 / 7529755066697.0;
            const double a32 = 11266239266428.0 / 11593286722821.0;
            double bi[16] = {0, 0, 0, 0, g, g, 0, 0, 2746238789719.0 / 10658868560708.0,
                             -640167445237.0 / 6845629431997.0, g, 0, a30, a31, a32, g};
            double be[16] = {0, 0, 0, 0, 1767732205903.0 / 2027836641118.0, 0, 0, 0,
                             5535828885825.0 / 10492691773637.0, 788022342437.0 / 10882634858940.0, 0, 0,
                             6485989280629.0 / 16251701735622.0, -4246266847089.0 / 9704473918619.0,
                             10755448449292.0 / 10357097424841.0, 0};
            double ba[4] = {a30, a31, a32, g};
            set_butcher_ass(s, 4, bi, be, ba, ba);
            s.wimp_lin0 = g; cf = 0.5;
            break;
        }
        case PZ_KC564: { /* :591-625 */
            double bi[36] = {0, 0, 0, 0, 0, 0,
                             0.25, 0.25, 0, 0, 0, 0,
                             8611.0 / 62500.0, -1743.0 / 31250.0, 0.25, 0, 0, 0,
                             5012029.0 / 34652500.0, -654441.0 / 2922500.0, 174375.0 / 388108.0, 0.25, 0, 0,
                             15267082809.0 / 155376265600.0, -71443401.0 / 120774400.0, 730878875.0 / 902184768.0,
                             2285395.0 / 8070912.0, 0.25, 0,
                             82889.0 / 524892.0, 0, 15625.0 / 83664.0, 69875.0 / 102672.0, -2260.0 / 8211.0, 0.25};
            double be[36] = {0, 0, 0, 0, 0, 0,
                             0.5, 0, 0, 0, 0, 0,
                             13861.0 / 62500.0, 6889.0 / 62500.0, 0, 0, 0, 0,
                             -116923316275.0 / 2393684061468.0, -2731218467317.0 / 15368042101831.0,
                             9408046702089.0 / 11113171139209.0, 0, 0, 0,
                             -451086348788.0 / 2902428689909.0, -2682348792572.0 / 7519795681897.0,
                             12662868775082.0 / 11960479115383.0, 3355817975965.0 / 11060851509271.0, 0, 0,
                             647845179188.0 / 3216320057751.0, 73281519250.0 / 8382639484533.0,
                             552539513391.0 / 3454668386233.0, 3354512671639.0 / 8306763924573.0,
                             4040.0 / 17871.0, 0};
            double ba[6] = {82889.0 / 524892.0, 0, 15625.0 / 83664.0, 69875.0 / 102672.0, -2260.0 / 8211.0, 0.25};
            set_butcher_ass(s, 6, bi, be, ba, ba);
            s.wimp_lin0 = 0.25; cf = 0.5;
            break;
        }
        case PZ_DBM453: { /* :626-651 */
            const double w = 0.32591194130117247;
            double bi[25] = {0, 0, 0, 0, 0,
                             -0.22284985318525410, w, 0, 0, 0,
                             -0.46801347074080545, 0.86349284225716961, w, 0, 0,
                             -0.46509906651927421, 0.81063103116959553, 0.61036726756832357, w, 0,
                             0.87795339639076675, -0.72692641526151547, 0.75204137157372720, -0.22898029400415088, w};
            double be[25] = {0, 0, 0, 0, 0,
                             0.10306208811591838, 0, 0, 0, 0,
                             -0.94124866143519894, 1.6626399742527356, 0, 0, 0,
                             -1.3670975201437765, 1.3815852911016873, 1.2673234025619065, 0, 0,
                             -0.81287582068772448, 0.81223739060505738, 0.90644429603699305, 0.094194134045674111, 0};
            double ba[5] = {0.87795339639076672, -0.72692641526151549, 0.7520413715737272, -0.22898029400415090, w};
            set_butcher_ass(s, 5, bi, be, ba, ba);
            s.wimp_lin0 = w; cf = 0.5;
            break;
        }
        case PZ_KC785: { /* :652-711 */
            const double w = 2.0 / 9.0;
            const double c2 = 2366667076620.0 / 8822750406821.0, c3 = -257962897183.0 / 4451812247028.0;
            const double c4 = -486229321650.0 / 11227943450093.0, c5 = 621307788657.0 / 4714163060173.0;
            const double c6 = 2036305566805.0 / 6583108094622.0;
            const double b2 = 3517720773327.0 / 20256071687669.0, b3 = 4569610470461.0 / 17934693873752.0;
            const double b4 = 2819471173109.0 / 11655438449929.0, b5 = 3296210113763.0 / 10722700128969.0;
            const double b6 = -1142099968913.0 / 5710983926999.0;
            double bi[64] = {0, 0, 0, 0, 0, 0, 0, 0,
                             w, w, 0, 0, 0, 0, 0, 0,
                             c2, c2, w, 0, 0, 0, 0, 0,
                             c3, c3, 128530224461.0 / 14379561246022.0, w, 0, 0, 0, 0,
                             c4, c4, -225633144460.0 / 6633558740617.0, 1741320951451.0 / 6824444397158.0, w, 0, 0, 0,
                             c5, c5, -125196015625.0 / 3866852212004.0, 940440206406.0 / 7593089888465.0,
                             961109811699.0 / 6734810228204.0, w, 0, 0,
                             c6, c6, -3039402635899.0 / 4450598839912.0, -1829510709469.0 / 31102090912115.0,
                             -286320471013.0 / 6931253422520.0, 8651533662697.0 / 9642993110008.0, w, 0,
                             0, 0, b2, b3, b4, b5, b6, w};
            const double e7 = 2193717860234.0 / 3570523412979.0;
            double be[64] = {0, 0, 0, 0, 0, 0, 0, 0,
                             4.0 / 9.0, 0, 0, 0, 0, 0, 0, 0,
                             1.0 / 9.0, 1183333538310.0 / 1827251437969.0, 0, 0, 0, 0, 0, 0,
                             895379019517.0 / 9750411845327.0, 477606656805.0 / 13473228687314.0,
                             -112564739183.0 / 9373365219272.0, 0, 0, 0, 0, 0,
                             -4458043123994.0 / 13015289567637.0, -2500665203865.0 / 9342069639922.0,
                             983347055801.0 / 8893519644487.0, 2185051477207.0 / 2551468980502.0, 0, 0, 0, 0,
                             -167316361917.0 / 17121522574472.0, 1605541814917.0 / 7619724128744.0,
                             991021770328.0 / 13052792161721.0, 2342280609577.0 / 11279663441611.0,
                             3012424348531.0 / 12792462456678.0, 0, 0, 0,
                             6680998715867.0 / 14310383562358.0, 5029118570809.0 / 3897454228471.0,
                             2415062538259.0 / 6382199904604.0, -3924368632305.0 / 6964820224454.0,
                             -4331110370267.0 / 15021686902756.0, -3944303808049.0 / 11994238218192.0, 0, 0,
                             e7, e7, 5952760925747.0 / 18750164281544.0, -4412967128996.0 / 6196664114337.0,
                             4151782504231.0 / 36106512998704.0, 572599549169.0 / 6265429158920.0,
                             -457874356192.0 / 11306498036315.0, 0};
            double ba[8] = {0, 0, b2, b3, b4, b5, b6, w};
            set_butcher_ass(s, 8, bi, be, ba, ba);
            s.wimp_lin0 = w; cf = 0.5;
            break;
        }
        default:
            return false;
        }
    }
    s.courfac = std::fabs(courfac_nml) >= 1.0e3 ? cf : courfac_nml;
    return true;
}

#define HIPCHK(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Adjoint buffers of a handle (pz_adjoint_forward/backward): the probed
// operators and scratch live as long as the handle; the snapshot blocks
// are sized for the largest window so far.
struct AdjCtx {
    PzAdjDev ad{};
    PzAdjDev *dad = nullptr; // device copy of ad
    PzDev *dlam = nullptr;   // device view of the adjoint variables
    char *lam = nullptr;     // adjoint variables (state layout)
    double *acc = nullptr;   // objective sum
    cd *W = nullptr, *grad = nullptr;
    cd *rrow = nullptr; // spectra by radius of the replayed states (long grids)
    // window: nseg checkpoints every ck steps, ck + 1 segment snapshots
    // and their ck forward views, the state at the end of the window
    char *ckpt = nullptr, *seg = nullptr, *fin = nullptr;
    PzDev *dviews = nullptr;
    int cap_ckpt = 0, cap_ck = 0, views_ck = 0; // views_ck: the ck the views were built for
    int objective = -1, K = 0, ck = 0, nseg = 0;
    bool ready = false; // a forward sweep waits for its backward sweep
    // copies of T_inv, X_inv, U_inv, PC_inv (steady, start-up)
    std::vector<double *> mats_b, mats_s;
    std::vector<size_t> mat_bytes;
    // the transposed steps run on s2 beside the replay of the previous
    // segment on the handle's stream; ev[0..1]: replay of buffer b done,
    // ev[2..3]: adjoint of buffer b done
    hipStream_t s2 = nullptr;
    hipEvent_t ev[4] = {};
    void release()
    {
        for (void *q : {(void *)ckpt, (void *)seg, (void *)fin, (void *)dviews})
            if (q) (void)hipFree(q);
        ckpt = seg = fin = nullptr;
        dviews = nullptr;
        cap_ckpt = cap_ck = 0;
    }
    ~AdjCtx()
    {
        release();
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        if (s2) (void)hipStreamDestroy(s2);
    }
};

} // namespace

struct pz_handle {
    pz_config c;
    int N = 0, nm = 0, nphi = 0, ncheb = 0, mmax = 0, minc = 1;
    std::vector<int> idx2m;
    double r_cmb = 0, r_icb = 0;
    bool l_heat = false, l_chem = false;
    std::vector<double> r, x, or1;
    std::vector<double> height, beta, drx;
    double rnorm = 0, cheb_fac = 0;
    Scheme sch;
    PzDev d{};
    std::vector<void *> allocs;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool started = false;
    bool use_graph = true;
    size_t lds_radial = 0, lds_mphase = 0, lds_build = 0, lds_start = 0, lds_inv = 0, lds_start_coll = 0, lds_fd = 0,
           lds_fd_loop = 0;
    size_t lds_gj = 0; // k_build_coll: blocked Gauss-Jordan panels
    bool mcoll_l2 = false; // collocation m-loop reads radial rows and tables through L2
    bool radial_big = false; // k_radial_loop_big for long azimuthal grids
    bool radial_inplace = false; // k_gather_r + k_radial_loop_inplace instead
    double dtcond_top = 0.0, dtcond_bot = 0.0; // dT_cond/dr at the walls (Nusselt numbers)
    char *arena = nullptr; // step state block (arena_layout)
    size_t arena_bytes = 0;
    AdjCtx *adj = nullptr; // adjoint buffers, made on first use
    int inv_waves = 16;
    int mthreads = PZ_MPHASE_THREADS; // threads of k_mphase / k_mcoll
    PzDev *ddev = nullptr;
    double2 *scratch = nullptr; // this device's copy of the staging array
    // restart state of pz_read_checkpoint, consumed by the next pz_start
    struct {
        bool pending = false;
        double time = 0.0, dt[8] = {};
        int ab1 = 0, bridge = 0;
    } rst;
    bool restarted = false; // the run started from a checkpoint
    std::vector<double> rintw; // rInt_R weights on this grid (time series)
    std::vector<double> spec;  // m-spectra of the last pz_time_series
    int bridge_default = 0; // l_bridge_step of a fresh start
    std::string err;

    template <class T> T *dalloc(size_t n)
    {
        void *p = nullptr;
        HIPCHK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
        HIPCHK(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)));
        allocs.push_back(p);
        return (T *)p;
    }
    template <class T> T *upload(const std::vector<T> &v)
    {
        T *p = dalloc<T>(v.size());
        HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return p;
    }
    ~pz_handle()
    {
        delete adj;
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void *p : allocs) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

static int fail(pz_handle *h, int code, const std::string &msg)
{
    if (h) h->err = msg;
    return code;
}

extern "C" void pz_config_default(pz_config *c)
{
    // namelists.f90:530-672
    std::memset(c, 0, sizeof(*c));
    c->n_r_max = 32;
    c->n_cheb_max = 32;
    c->m_max = 32;
    c->minc = 1;
    c->time_scheme = PZ_CNAB2;
    c->l_cheb_coll = 1;
    c->l_direct_solve = 1;
    c->l_coriolis_imp = 1;
    c->l_buo_imp = 1;
    c->rerror_fac = 500.0;
    c->alpha = 0.5;
    c->dtMax = 1.0e-5;
    c->courfac = 1.0e3;
    c->dt_fac = 2.0;
    c->ra = 1.0e5;
    c->ek = 1.0e-3;
    c->pr = 1.0;
    c->radratio = 0.35;
    c->g1 = 1.0;
    c->ktopt = c->kbott = c->ktopxi = c->kbotxi = 1;
    c->ktopv = c->kbotv = 2;
    c->tcond_fac = c->xicond_fac = 1.0;
    c->fd_stretch = 0.3; // namelists.f90:544-545
    c->fd_ratio = 0.1;
    c->l_finite_diff = 0;
}

extern "C" const char *pz_last_error(const pz_handle *h) { return h ? h->err.c_str() : "null handle"; }

// ---------------------------------------------------------------------------
// set-up helpers
// ---------------------------------------------------------------------------
static void conducting_state(const pz_handle *h, double cond_fac, int ktop, int kbot, const double *t_top,
                             const double *t_bot, std::vector<double> &tc, std::vector<double> &dtc)
{
    // get_conducting_state, 2-D branch (radial.f90:284-397)
    const double eps10 = 10.0 * 2.220446049250313e-16;
    double f_bot = 0, f_top = 0, epsc0 = 0.0;
    for (int n = 0; n < PZ_N_T_BOUNDS; n++) {
        if ((int)t_bot[3 * n] == 0 && std::fabs(t_bot[3 * n + 1]) > eps10) f_bot = t_bot[3 * n + 1];
        if ((int)t_top[3 * n] == 0 && std::fabs(t_top[3 * n + 1]) > eps10) f_top = t_top[3 * n + 1];
    }
    const double rc = h->r_cmb, ri = h->r_icb, rr = h->c.radratio;
    tc.assign(h->N, 0.0);
    dtc.assign(h->N, 0.0);
    for (int n = 0; n < h->N; n++) {
        double r = h->r[n], o1 = h->or1[n];
        if (ktop == 1 && kbot == 1) {
            tc[n] = cond_fac * (std::log(r / rc) / std::log(rr));
            dtc[n] = cond_fac * (o1 / std::log(rr));
        } else if (ktop == 2 && kbot == 1) {
            tc[n] = cond_fac * (f_bot + f_top * rc * std::log(r / ri) + 0.25 * epsc0 * (ri * ri - r * r + 2.0 * rc * rc * std::log(r / ri)));
            dtc[n] = cond_fac * (f_top * rc * o1 + 0.5 * epsc0 * (rc * rc * o1 - r));
        } else if (ktop == 1 && kbot == 2) {
            tc[n] = cond_fac * (f_bot * ri * std::log(r / rc) + 0.25 * epsc0 * (rc * rc - r * r + 2.0 * ri * ri * std::log(r / rc)) + f_top);
            dtc[n] = cond_fac * (f_bot * ri * o1 + 0.5 * epsc0 * (ri * ri * o1 - r));
        } else {
            double e0 = -2.0 * (rc * f_top - ri * f_bot) / (rc * rc - ri * ri);
            if (std::fabs(e0) < eps10) e0 = 0.0;
            tc[n] = cond_fac * (0.5 * e0 * (rc * rc * std::log(r / rc) - 0.5 * (r * r - rc * rc)) + rc * f_top * std::log(r / rc));
            dtc[n] = cond_fac * (f_top * rc * o1 + 0.5 * e0 * (rc * rc * o1 - r));
        }
    }
}

static std::vector<double2> bc_pattern(const pz_handle *h, const double *tb, const double *tt, bool top)
{
    // horizontal.f90:149-173
    std::vector<double2> v(h->nm, make_double2(0, 0));
    for (int i = 0; i < h->nm; i++) {
        int m = h->idx2m[i];
        for (int n = 0; n < PZ_N_T_BOUNDS; n++) {
            const double *t = top ? tt : tb;
            double tr = t[3 * n + 1], ti = t[3 * n + 2];
            if ((int)t[3 * n] == m && (tr != 0 || ti != 0)) {
                if (m == 0) ti = 0;
                v[i] = make_double2(tr, ti);
            }
        }
    }
    return v;
}

// row-major mat-vec band [N][2K+1], column-offset order, x rnorm
static std::vector<double> rhs_band(const pz_handle *h, int K, int nbounds, const std::vector<ld> *st_rows)
{
    const int N = h->N, W = 2 * K + 1;
    std::vector<double> b((size_t)N * W, 0.0);
    for (int i = nbounds; i < N; i++)
        for (int d = -K; d <= K; d++) {
            int col = i + d;
            if (col < 0 || col >= N) continue;
            b[(size_t)i * W + K + d] = h->rnorm * (double)st_rows[i][K - d];
        }
    return b;
}

static void build_operators(pz_handle *h)
{
    const int N = h->N, nm = h->nm;
    pzh::OpLib L;
    L.build(N, (ld)h->r_cmb, (ld)h->r_icb);
    auto rows = [&](int K, std::vector<const pzh::ExtTab *> tabs, std::vector<ld> coef) {
        std::vector<std::vector<ld>> out(N);
        for (int i = 0; i < N; i++) out[i] = pzh::stencil(i, K, tabs, coef);
        return out;
    };
    auto rnd = [](const std::vector<std::vector<ld>> &rw, std::vector<double> &dst, size_t base) {
        size_t W = rw[0].size();
        for (size_t i = 0; i < rw.size(); i++)
            for (size_t k = 0; k < W; k++) dst[base + i * W + k] = (double)rw[i][k];
    };
    // m-independent RHS bands
    {
        auto s = rows(3, {&L.I2R1}, {1});
        h->d.T_RHSE1 = h->upload(rhs_band(h, 3, 2, s.data()));
        s = rows(4, {&L.I2R2}, {1});
        h->d.T_RHSE2 = h->upload(rhs_band(h, 4, 2, s.data()));
        h->d.P_RHSE1 = h->upload(rhs_band(h, 4, 2, s.data()));
        s = rows(8, {&L.I4R4}, {1});
        h->d.P_RHSE2 = h->upload(rhs_band(h, 8, 4, s.data()));
        s = rows(7, {&L.I4R3}, {1});
        h->d.P_RHSE3 = h->upload(rhs_band(h, 7, 4, s.data()));
    }
    std::vector<double> TRHSI((size_t)nm * N * 5, 0.0), PRHSI((size_t)nm * N * 17, 0.0);
    std::vector<double2> PRHSIL((size_t)nm * N * 17, make_double2(0, 0));
    std::vector<double> TS0((size_t)nm * N * 9), TS1((size_t)nm * N * 9);
    std::vector<double> PS0((size_t)nm * N * 17, 0.0), PS1((size_t)nm * N * 17, 0.0), PS2((size_t)nm * N * 17, 0.0);
    auto i2r2_4 = rows(4, {&L.I2R2}, {1});
    auto i4r4_8 = rows(8, {&L.I4R4}, {1});
    auto r2_4 = rows(4, {&L.R2}, {1}), i1r1_4 = rows(4, {&L.I1R1}, {1});
    auto r2_2 = rows(2, {&L.R2}, {1}), i1r1_2 = rows(2, {&L.I1R1}, {1});
    for (int mi = 0; mi < nm; mi++) {
        const int m = h->idx2m[mi];
        const ld M2 = (ld)m * (ld)m;
        const double nu = h->d.ViscFac * 1.0; // hdif_V = 1
        // T / xi (update_temp_integ.f90:538-846)
        auto lapl2 = rows(2, {&L.L2_0, &L.L2_2}, {1, M2});
        std::vector<double> band = rhs_band(h, 2, 2, lapl2.data());
        std::copy(band.begin(), band.end(), TRHSI.begin() + (size_t)mi * N * 5);
        rnd(i2r2_4, TS0, (size_t)mi * N * 9);
        rnd(rows(4, {&L.L2_0, &L.L2_2}, {1, M2}), TS1, (size_t)mi * N * 9);
        // psi (update_psi_integ_smat.f90:1117-1764), rotating sphere
        const size_t pb = (size_t)mi * N * 17;
        if (m == 0) {
            band = rhs_band(h, 4, 2, i2r2_4.data());
            std::copy(band.begin(), band.end(), PRHSI.begin() + pb);
            // nu*(rmult2 - 3 intcheb1rmult1), K = 2
            for (int i = 2; i < N; i++)
                for (int d = -2; d <= 2; d++) {
                    int col = i + d;
                    if (col < 0 || col >= N) continue;
                    double sc = nu * ((double)r2_2[i][2 - d] - 3.0 * (double)i1r1_2[i][2 - d]);
                    PRHSIL[pb + (size_t)i * 5 + 2 + d] = make_double2(h->rnorm * sc, 0.0);
                }
            rnd(i2r2_4, PS0, pb);
            for (int i = 0; i < N; i++)
                for (int k = 0; k < 9; k++) PS1[pb + (size_t)i * 9 + k] = (double)r2_4[i][k] - 3.0 * (double)i1r1_4[i][k];
        } else {
            auto lr = rows(8, {&L.LR_0, &L.LR_2}, {1, M2});
            auto lr2 = rows(8, {&L.LR2_0, &L.LR2_2, &L.LR2_4}, {1, M2, M2 * M2});
            band = rhs_band(h, 8, 4, lr.data());
            std::copy(band.begin(), band.end(), PRHSI.begin() + pb);
            const double corf = h->c.l_coriolis_imp ? -h->d.CorFac * (double)m : 0.0;
            for (int i = 4; i < N; i++)
                for (int d = -8; d <= 8; d++) {
                    int col = i + d;
                    if (col < 0 || col >= N) continue;
                    double sc = nu * (double)lr2[i][8 - d];
                    double cs = corf * (double)i4r4_8[i][8 - d];
                    PRHSIL[pb + (size_t)i * 17 + 8 + d] = make_double2(h->rnorm * sc, h->rnorm * cs);
                }
            rnd(lr, PS0, pb);
            rnd(lr2, PS1, pb);
            rnd(i4r4_8, PS2, pb);
        }
    }
    h->d.T_RHSI = h->upload(TRHSI);
    h->d.P_RHSI = h->upload(PRHSI);
    h->d.P_RHSIL = h->upload(PRHSIL);
    h->d.T_S0 = h->upload(TS0);
    h->d.T_S1 = h->upload(TS1);
    h->d.X_S1 = h->d.T_S1;
    h->d.P_S0 = h->upload(PS0);
    h->d.P_S1 = h->upload(PS1);
    h->d.P_S2 = h->upload(PS2);
}

static std::vector<int> factor_radices(int n)
{
    std::vector<int> f;
    while (n % 8 == 0) { f.push_back(8); n /= 8; }
    while (n % 4 == 0) { f.push_back(4); n /= 4; }
    while (n % 2 == 0) { f.push_back(2); n /= 2; }
    while (n % 3 == 0) { f.push_back(3); n /= 3; }
    for (int p = 5; n > 1; p += 2)
        while (n % p == 0) { f.push_back(p); n /= p; }
    return f;
}

// Fornberg's finite-difference weights (Math. Comp. 51 (1988) 699-706), as
// in populate_fd_weights (finite_differences.f90:501-546): c[j * (m + 1) + k]
// is the weight of x[j] in the k-th derivative at z.
static void fd_weights(double z, const double *x, int nd, int m, double *c)
{
    const int w = m + 1;
    double c1 = 1.0, c4 = x[0] - z;
    for (int i = 0; i < (nd + 1) * w; i++) c[i] = 0.0;
    c[0] = 1.0;
    for (int i = 1; i <= nd; i++) {
        const int mn = std::min(i, m);
        double c2 = 1.0, c5 = c4;
        c4 = x[i] - z;
        for (int j = 0; j < i; j++) {
            const double c3 = x[i] - x[j];
            c2 *= c3;
            if (j == i - 1) {
                for (int k = mn; k >= 1; k--)
                    c[i * w + k] = c1 * ((double)k * c[(i - 1) * w + k - 1] - c5 * c[(i - 1) * w + k]) / c2;
                c[i * w] = -c1 * c5 * c[(i - 1) * w] / c2;
            }
            for (int k = mn; k >= 1; k--) c[j * w + k] = (c4 * c[j * w + k] - (double)k * c[j * w + k - 1]) / c3;
            c[j * w] = c4 * c[j * w] / c3;
        }
        c1 = c2;
    }
}

// get_FD_grid (finite_differences.f90:87-188) for an annulus, r[0] = r_cmb
static std::vector<double> fd_grid(int N, double rcmb, double ricb, double ratio1, double ratio2)
{
    std::vector<double> r(N);
    r[0] = rcmb;
    double drb;
    if (ratio2 == 1.0) {
        drb = 1.0 / ((double)N - 1.0);
        for (int k = 1; k < N; k++) r[k] = r[k - 1] - drb;
    } else {
        const int nbp = (int)((double)(N - 1) / (2.0 * (1.0 + ratio1)));
        const int nbulk = N - 1 - 2 * nbp;
        if (nbp < 1) throw std::invalid_argument("fd_stretch leaves no boundary points");
        const double dra = std::exp(std::log(ratio2) / (double)nbp);
        drb = 1.0;
        for (int i = 0; i < nbp; i++) drb *= dra;
        drb = 1.0 / ((double)nbulk + 2.0 * dra * ((1.0 - drb) / (1.0 - dra)));
        for (int i = 0; i < nbp; i++) drb *= dra;
        for (int k = 1; k <= nbp; k++) { r[k] = r[k - 1] - drb; drb /= dra; }
        for (int k = 1; k <= nbulk; k++) r[k + nbp] = r[k + nbp - 1] - drb;
        for (int k = 1; k <= nbp; k++) { drb *= dra; r[k + nbp + nbulk] = r[k + nbp + nbulk - 1] - drb; }
    }
    if (r[N - 1] != ricb) {
        if (std::fabs(r[N - 1] - ricb) > drb) throw std::invalid_argument("wrong internal radius of the FD grid");
        r[N - 1] = ricb;
    }
    return r;
}

// get_FD_coeffs, orders 2 and 2 (finite_differences.f90:190-446): the
// stencils of the step, laid out as PzDev::fdc
static std::vector<double> fd_coeffs(const std::vector<double> &r)
{
    const int N = (int)r.size();
    std::vector<double> c(16 * (size_t)N + 6, 0.0);
    double x[5], w[25];
    for (int k = 0; k < N; k++) {
        double *d1 = &c[3 * (size_t)k], *d2 = &c[3 * (size_t)N + 3 * k];
        if (k == 0 || k == N - 1) {
            const double drl = k == 0 ? r[1] - r[0] : r[N - 1] - r[N - 2];
            d1[0] = -0.5 / drl; d1[1] = 0.0; d1[2] = 0.5 / drl;
            d2[0] = 1.0 / drl / drl; d2[1] = -2.0 / drl / drl; d2[2] = 1.0 / drl / drl;
        } else {
            for (int od = 0; od < 3; od++) x[od] = r[k - 1 + od] - r[k];
            fd_weights(0.0, x, 2, 2, w);
            for (int od = 0; od < 3; od++) { d1[od] = w[od * 3 + 1]; d2[od] = w[od * 3 + 2]; }
        }
    }
    for (int k = 1; k < N - 1; k++) {
        for (int od = 0; od < 5; od++) {
            if (k == 1 && od == 0) x[od] = 2.0 * (r[0] - r[1]);
            else if (k == N - 2 && od == 4) x[od] = 2.0 * (r[N - 1] - r[N - 2]);
            else x[od] = r[k - 2 + od] - r[k];
        }
        fd_weights(0.0, x, 4, 4, w);
        for (int od = 0; od < 5; od++) {
            c[6 * (size_t)N + 5 * k + od] = w[od * 5 + 3];
            c[11 * (size_t)N + 5 * k + od] = w[od * 5 + 4];
        }
    }
    for (int od = 0; od < 3; od++) x[od] = r[od] - r[0];
    fd_weights(0.0, x, 2, 2, w);
    for (int od = 0; od < 3; od++) c[16 * (size_t)N + od] = w[od * 3 + 1];
    for (int od = 0; od < 3; od++) x[od] = r[N - 1 - od] - r[N - 1];
    fd_weights(0.0, x, 2, 2, w);
    for (int od = 0; od < 3; od++) c[16 * (size_t)N + 3 + od] = w[od * 3 + 1];
    return c;
}

// Static LDS of the collocation rebuild kernels (the larger of the single-run
// and batched variants), read from the compiled kernels once per process.
static size_t gj_static_lds()
{
    static size_t bytes = 0;
    if (!bytes) {
        hipFuncAttributes a, b;
        HIPCHK(hipFuncGetAttributes(&a, (const void *)k_build_coll));
        HIPCHK(hipFuncGetAttributes(&b, (const void *)k_build_coll_batch));
        bytes = std::max<size_t>({a.sharedSizeBytes, b.sharedSizeBytes, 1});
    }
    return bytes;
}

// Lets every kernel with dynamic LDS use the whole per-workgroup budget, once
// per process: a per-run value would shrink the limit for a larger run
// created earlier (runs of different sizes share the kernels).
static void allow_full_lds(int lim)
{
    static int done = 0;
    if (done >= lim) return;
    const void *kernels[] = {(const void *)k_build_lhs, (const void *)k_mphase, (const void *)k_radial_loop,
                             (const void *)k_start, (const void *)k_build_inv, (const void *)k_mcoll,
                             (const void *)k_start_coll, (const void *)k_build_lhs_batch,
                             (const void *)k_mphase_batch, (const void *)k_radial_loop_batch,
                             (const void *)k_build_inv_batch, (const void *)k_mcoll_batch<true>, (const void *)k_mcoll_batch<false>, (const void *)k_fd_build,
                             (const void *)k_fd_mloop, (const void *)k_fd_start, (const void *)k_radial_loop_big,
                             (const void *)k_radial_loop_inplace,
                             (const void *)k_adj_probe, (const void *)k_mcoll_adj, (const void *)k_radial_adj,
                             (const void *)k_start_coll_adj, (const void *)k_assemble_coll,
                             (const void *)k_fd_mloop_adj, (const void *)k_fd_start_adj,
                             (const void *)k_radial_adj_inplace,
                             (const void *)k_assemble_fd, (const void *)k_build_coll,
                             (const void *)k_build_coll_batch, (const void *)k_mcoll_l2};
    for (const void *k : kernels) {
        hipFuncAttributes a;
        HIPCHK(hipFuncGetAttributes(&a, k));
        // the dynamic part may use what the kernel's static LDS leaves
        HIPCHK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, lim - (int)a.sharedSizeBytes));
    }
    done = lim;
}

static void setup(pz_handle *h)
{
    const pz_config &c = h->c;
    const int N = h->N = c.n_r_max;
    h->mmax = c.m_max;
    h->minc = c.minc;
    h->nm = c.m_max / c.minc + 1;             // truncation.f90:28
    h->nphi = 3 * c.m_max / c.minc;           // truncation.f90:30-31
    h->ncheb = c.n_cheb_max;
    h->idx2m.resize(h->nm);
    for (int i = 0; i < h->nm; i++) h->idx2m[i] = i * c.minc;
    if (!make_scheme(c.time_scheme, c.courfac, h->sch)) throw std::invalid_argument("unsupported time scheme");
    h->l_heat = c.ra != 0.0;
    h->l_chem = c.raxi != 0.0;
    h->r_cmb = 1.0 / (1.0 - c.radratio);
    h->r_icb = h->r_cmb - 1.0;
    PzDev &d = h->d;
    d.N = N;
    d.nm = h->nm;
    d.nphi = h->nphi;
    d.ncheb = c.n_cheb_max;
    d.L = 2 * N - 2;
    d.family = h->sch.family;
    d.scheme = h->sch.id;
    // a fresh start sets l_bridge_step (namelists.f90:364-365): the first
    // nold-1 steps of a multistep scheme other than CNAB2 are CNAB2 steps
    d.bridge = h->sch.family == 0 && h->sch.id != PZ_CNAB2;
    h->bridge_default = d.bridge;
    d.nstages = h->sch.nstages;
    d.nexp = h->sch.nexp;
    d.nimp = h->sch.nimp;
    d.nold = h->sch.nold;
    for (int k = 0; k < 8; k++) { d.l_exp_calc[k] = h->sch.l_exp_calc[k]; d.l_imp_calc_rhs[k] = h->sch.l_imp_calc_rhs[k]; }
    std::memcpy(d.bimp0, h->sch.bimp0, sizeof(d.bimp0));
    std::memcpy(d.bexp0, h->sch.bexp0, sizeof(d.bexp0));
    d.l_assembly = h->sch.l_assembly;
    std::memcpy(d.bass_imp0, h->sch.bass_imp0, sizeof(d.bass_imp0));
    std::memcpy(d.bass_exp0, h->sch.bass_exp0, sizeof(d.bass_exp0));
    d.wimp_lin0 = h->sch.wimp_lin0;
    d.courfac = h->sch.courfac;
    d.l_heat = h->l_heat;
    d.l_chem = h->l_chem;
    // namelists.f90:448-457: the influence-matrix method treats Coriolis
    // explicitly
    d.l_coriolis_imp = (!c.l_direct_solve && !c.l_non_rot) ? 0 : c.l_coriolis_imp;
    h->c.l_coriolis_imp = d.l_coriolis_imp;
    d.l_buo_imp = c.l_buo_imp;
    d.l_ek_pump = c.l_ek_pump;
    d.ktopt = c.ktopt; d.kbott = c.kbott; d.ktopxi = c.ktopxi; d.kbotxi = c.kbotxi;
    d.ktopv = c.ktopv; d.kbotv = c.kbotv;
    // prefactors (namelists.f90:471-526)
    if (c.time_scale_rot) {
        d.CorFac = 2.0;
        d.BuoFac = h->l_heat ? c.ra * c.ek * c.ek / c.pr : 0.0;
        d.TdiffFac = h->l_heat ? c.ek / c.pr : 0.0;
        d.ChemFac = h->l_chem ? c.raxi * c.ek * c.ek / c.sc : 0.0;
        d.XidiffFac = h->l_chem ? c.ek / c.sc : 0.0;
        d.ViscFac = c.ek;
    } else {
        d.CorFac = 2.0 / c.ek;
        d.BuoFac = h->l_heat ? c.ra / c.pr : 0.0;
        d.TdiffFac = h->l_heat ? 1.0 / c.pr : 0.0;
        d.ChemFac = h->l_chem ? c.raxi / c.sc : 0.0;
        d.XidiffFac = h->l_chem ? 1.0 / c.sc : 0.0;
        d.ViscFac = 1.0;
    }
    if (c.l_non_rot) {
        // namelists.f90:473-489
        d.CorFac = 0.0;
        d.BuoFac = h->l_heat ? c.ra / c.pr : 0.0;
        d.TdiffFac = h->l_heat ? 1.0 / c.pr : 0.0;
        d.ChemFac = h->l_chem ? c.raxi / c.sc : 0.0;
        d.XidiffFac = h->l_chem ? 1.0 / c.sc : 0.0;
        d.ViscFac = 1.0;
    }
    d.tadvz = c.l_temp_advz ? 1.0 : 0.0;
    d.r_cmb = h->r_cmb;
    d.ek = c.l_non_rot ? -1.0 : c.ek; // namelists.f90:312-315: ek = -1 flags a non-rotating run
    d.alpha = c.alpha;
    d.dtMax = c.dtMax;
    d.dtMin = c.dtMax / 1.0e6; // pre_calculations.f90:30
    d.dt_fac = c.dt_fac;
    // Chebyshev constants (chebyshev.f90:55-60, dct_fftw.f90:124)
    h->rnorm = d.rnorm = std::sqrt(2.0 / (double)(N - 1));
    h->cheb_fac = d.cheb_fac = std::sqrt(0.5 / (double)(N - 1));
    d.icheb_fac = 1.0 / d.cheb_fac;
    d.boundary_fac = 0.5;
    d.thr = c.l_rerror_fix ? c.rerror_fac * 2.220446049250313e-16 : 0.0;
    // grid (chebyshev.f90:188-240; radial.f90:115-282)
    const double pi = 4.0 * std::atan(1.0);
    const double bma = 0.5 * (h->r_cmb - h->r_icb), bpa = 0.5 * (h->r_icb + h->r_cmb), rc = h->r_cmb;
    std::vector<double> x(N), r(N), drx(N), ddrx(N, 0.0), or1(N), or2(N), rgrav(N), beta(N, 0.0), ekpump(N, 0.0);
    std::vector<double> ekp_up(N), ekp_us(N), ekp_dusdp(N), height(N, 0.0), oheight(N, 0.0), delxr2(N), delxh2(N);
    for (int k = 0; k < N; k++) {
        x[k] = std::cos(pi * (double)k / (double)(N - 1));
        r[k] = bma * x[k] + bpa;
        drx[k] = 2.0 / (h->r_cmb - h->r_icb);
    }
    d.fd = c.l_finite_diff ? 1 : 0;
    d.fd_pattern = c.l_finite_diff && c.l_fd_bc_pattern ? 1 : 0;
    if (d.fd) r = fd_grid(N, h->r_cmb, h->r_icb, c.fd_stretch, c.fd_ratio);
    for (int k = 0; k < N; k++) {
        or1[k] = 1.0 / r[k];
        or2[k] = or1[k] * or1[k];
        rgrav[k] = c.g0 + c.g1 * r[k] / rc + c.g2 * rc * rc * or2[k];
    }
    h->r = r;
    h->x = x;
    h->or1 = or1;
    h->drx = drx;
    const double ek_pump_fac = (c.l_ek_pump && !c.l_non_rot) ? 1.0 : 0.0;
    for (int n = 1; n < N; n++) {
        double s2 = rc * rc - r[n] * r[n];
        height[n] = 2.0 * std::sqrt(s2);
        oheight[n] = 0.5 / std::sqrt(s2);
        beta[n] = -r[n] / s2;
        ekpump[n] = 0.5 * ek_pump_fac * std::sqrt(c.ek * rc) / std::pow(s2, 0.75);
    }
    std::vector<double> dbeta(N, 0.0), d2beta(N, 0.0), d3beta(N, 0.0), or3(N); // radial.f90:212-226
    for (int n = 1; n < N; n++) {
        double s2 = rc * rc - r[n] * r[n];
        dbeta[n] = -(rc * rc + r[n] * r[n]) / (s2 * s2);
        d2beta[n] = -128.0 * r[n] * (3.0 * rc * rc + r[n] * r[n]) / std::pow(height[n], 6);
        d3beta[n] = -1536.0 * (std::pow(rc, 4) + std::pow(r[n], 4) + 6.0 * r[n] * r[n] * rc * rc) / std::pow(height[n], 8);
    }
    for (int n = 0; n < N; n++) or3[n] = or2[n] * or1[n];
    for (int n = 0; n < N; n++) {
        ekp_up[n] = -0.5 * beta[n];
        ekp_us[n] = -5.0 * rc * beta[n] * oheight[n];
        ekp_dusdp[n] = beta[n];
    }
    if (c.l_non_rot)
        // radial.f90:197-207
        for (int n = 0; n < N; n++) {
            beta[n] = dbeta[n] = d2beta[n] = d3beta[n] = 0.0;
            ekpump[n] = ekp_up[n] = ekp_us[n] = ekp_dusdp[n] = 0.0;
            height[n] = oheight[n] = 1.0;
        }
    h->height = height;
    h->beta = beta;
    const int mR = c.m_max;
    delxh2[0] = std::pow(2.0 * pi / 3.0 / mR * rc, 2);
    delxh2[N - 1] = std::pow(2.0 * pi / 3.0 / mR * h->r_icb, 2);
    delxr2[0] = std::pow(r[0] - r[1], 2);
    delxr2[N - 1] = std::pow(r[N - 2] - r[N - 1], 2);
    for (int n = 1; n < N - 1; n++) {
        delxh2[n] = std::pow(2.0 * pi / 3.0 / mR * r[n], 2);
        double dm = std::min(r[n - 1] - r[n], r[n] - r[n + 1]);
        delxr2[n] = dm * dm;
    }
    std::vector<double> tcond(N, 0.0), dtcond(N, 0.0), xicond(N, 0.0), dxicond(N, 0.0);
    if (h->l_heat) conducting_state(h, c.tcond_fac, c.ktopt, c.kbott, c.t_top, c.t_bot, tcond, dtcond);
    h->dtcond_top = dtcond[0];
    h->dtcond_bot = dtcond[N - 1];
    if (h->l_chem) conducting_state(h, c.xicond_fac, c.ktopxi, c.kbotxi, c.xi_top, c.xi_bot, xicond, dxicond);
    // transform tables
    const int L = 2 * N - 2;
    std::vector<double> cosT(L), sinT(L);
    const ld lpi = 3.14159265358979323846264338327950288L;
    for (int q = 0; q < L; q++) {
        cosT[q] = (double)cosl(lpi * (ld)q / (ld)(N - 1));
        sinT[q] = (double)sinl(lpi * (ld)q / (ld)(N - 1));
    }
    d.cosT = h->upload(cosT);
    d.sinT = h->upload(sinT);
    // bulk factors of get_dr_fft / get_ddr_fft at the interior nodes:
    // 1/(sqrt(1-x^2) nrm), x/((1-x^2)^1.5 nrm), 1/((1-x^2) nrm)
    std::vector<double> drb(3 * N, 0.0);
    {
        const ld nrm = (ld)(2 * N - 2);
        for (int j = 1; j < N - 1; j++) {
            const double xj = x[j];
            const ld sj = (ld)(1.0 - xj * xj);
            drb[j] = (double)(1.0L / (sqrtl(sj) * nrm));
            drb[N + j] = (double)((ld)xj / (powl(sj, 1.5L) * nrm));
            drb[2 * N + j] = (double)(1.0L / (sj * nrm));
        }
    }
    {
        // all radial functions in one block of PZ_NRADC rows of n_r values
        // (row order of set_radial_pointers), so that a kernel can stage
        // them into LDS with one copy
        const std::vector<double> *rows[18] = {&x, &r, &drx, &ddrx, &or1, &or2, &rgrav, &beta, &ekpump,
                                               &ekp_up, &ekp_us, &ekp_dusdp, &delxr2, &delxh2, &tcond,
                                               &dtcond, &xicond, &dxicond};
        std::vector<double> rc((size_t)PZ_NRADC * N, 0.0);
        for (int i = 0; i < 18; i++) std::copy(rows[i]->begin(), rows[i]->end(), rc.begin() + (size_t)i * N);
        std::copy(drb.begin(), drb.end(), rc.begin() + (size_t)18 * N);
        std::copy(dbeta.begin(), dbeta.end(), rc.begin() + (size_t)21 * N);
        std::copy(or3.begin(), or3.end(), rc.begin() + (size_t)22 * N);
        std::copy(d2beta.begin(), d2beta.end(), rc.begin() + (size_t)23 * N);
        std::copy(d3beta.begin(), d3beta.end(), rc.begin() + (size_t)24 * N);
        d.radc = h->upload(rc);
        set_radial_pointers(d, d.radc, (size_t)N);
    }
    std::vector<double> tau(6 * N);
    for (int n = 0; n < N; n++) {
        // chebyshev.f90:362-381
        double dn2 = (double)n * (double)n, monen = (n % 2 == 0) ? 1.0 : -1.0;
        tau[n] = 1.0;
        tau[N + n] = monen;
        tau[2 * N + n] = 2.0 * dn2;
        tau[3 * N + n] = 2.0 * dn2 * (-1.0) * monen;
        tau[4 * N + n] = 4.0 * (1.0 / 3.0) * dn2 * (dn2 - 1.0);
        tau[5 * N + n] = 4.0 * (1.0 / 3.0) * dn2 * (dn2 - 1.0) * monen;
    }
    d.tau = h->upload(tau);
    std::vector<double2> tw(h->nphi);
    for (int k = 0; k < h->nphi; k++) {
        ld ang = 2.0L * lpi * (ld)k / (ld)h->nphi;
        tw[k] = make_double2((double)cosl(ang), -(double)sinl(ang));
    }
    d.twid = h->upload(tw);
    std::vector<int> rad = factor_radices(h->nphi);
    if ((int)rad.size() > PZ_MAX_RADICES || *std::max_element(rad.begin(), rad.end()) > 64)
        throw std::invalid_argument("n_phi has a prime factor above 64");
    d.nrad = (int)rad.size();
    for (size_t i = 0; i < rad.size(); i++) d.radices[i] = rad[i];
    // Chebyshev transforms by FFTs of length L unless L has a prime factor
    // above 64 (then the direct cosine sums are used)
    std::vector<int> radL = factor_radices(2 * N - 2);
    d.nradL = (radL.size() <= PZ_MAX_RADICES && *std::max_element(radL.begin(), radL.end()) <= 64) ? (int)radL.size() : 0;
    for (int i = 0; i < d.nradL; i++) d.radicesL[i] = radL[i];
    d.idx2m = h->upload(h->idx2m);
    std::vector<double> ones(h->nm, 1.0); // no hyperdiffusion (horizontal.f90:126-131)
    d.hdifT = h->upload(ones);
    d.hdifXi = h->upload(ones);
    d.hdifV = h->upload(ones);
    d.topt = h->upload(bc_pattern(h, c.t_bot, c.t_top, true));
    d.bott = h->upload(bc_pattern(h, c.t_bot, c.t_top, false));
    // the composition pattern is read from t_bot/t_top (horizontal.f90:176-198)
    d.topxi = h->upload(bc_pattern(h, c.t_bot, c.t_top, true));
    d.botxi = h->upload(bc_pattern(h, c.t_bot, c.t_top, false));
    // fields and time arrays
    const size_t F = (size_t)h->nm * N;
    // the state a step carries to the next (fields, time arrays, scheme
    // state) lives in one block, so that a snapshot is one copy
    const Scheme &s = h->sch;
    h->arena_bytes = arena_layout(d, nullptr, F, s.nold, s.nimp, s.nexp);
    h->arena = h->dalloc<char>(h->arena_bytes);
    arena_layout(d, h->arena, F, s.nold, s.nimp, s.nexp);
    d.dVsT = h->dalloc<cd>(F); d.dVsXi = h->dalloc<cd>(F); d.dVsOm = h->dalloc<cd>(F);
    d.dtr = h->dalloc<double>(N);
    d.dth = h->dalloc<double>(N);
    // factorised matrices
    const int nm = h->nm, nb2 = N - 2;
    d.T_A1 = h->dalloc<double>(nm * 4); d.T_A2 = h->dalloc<double>((size_t)nm * 2 * nb2);
    d.T_A3 = h->dalloc<double>((size_t)nm * 2 * nb2); d.T_A4 = h->dalloc<double>((size_t)nm * PZ_T_LD * nb2);
    d.T_p1 = h->dalloc<int>(nm * 2); d.T_p4 = h->dalloc<int>((size_t)nm * nb2); d.T_fac = h->dalloc<double>(F);
    d.T_dinv = h->dalloc<double>((size_t)nm * nb2); d.X_dinv = h->dalloc<double>((size_t)nm * nb2);
    d.P_dinv = h->dalloc<cd>((size_t)nm * N);
    d.X_A1 = h->dalloc<double>(nm * 4); d.X_A2 = h->dalloc<double>((size_t)nm * 2 * nb2);
    d.X_A3 = h->dalloc<double>((size_t)nm * 2 * nb2); d.X_A4 = h->dalloc<double>((size_t)nm * PZ_T_LD * nb2);
    d.X_p1 = h->dalloc<int>(nm * 2); d.X_p4 = h->dalloc<int>((size_t)nm * nb2); d.X_fac = h->dalloc<double>(F);
    d.P_A1 = h->dalloc<cd>(nm * 16); d.P_A2 = h->dalloc<cd>((size_t)nm * PZ_P_NT * N);
    d.P_A3 = h->dalloc<cd>((size_t)nm * PZ_P_NT * N); d.P_A4 = h->dalloc<cd>((size_t)nm * PZ_P_LD * N);
    d.P_p1 = h->dalloc<int>(nm * 4); d.P_p4 = h->dalloc<int>((size_t)nm * N); d.P_fac = h->dalloc<double>(F);
    d.info = h->dalloc<int>(4 * nm);
    d.ldi = (N + 1) & ~1;
    d.T_inv = h->dalloc<double>((size_t)nm * N * d.ldi);
    d.X_inv = h->dalloc<double>((size_t)nm * N * d.ldi);
    d.P_inv = h->dalloc<cd>((size_t)nm * N * d.ldi);
    d.coll = (c.l_cheb_coll && !d.fd) ? 1 : 0;
    if (d.fd) {
        d.fdc = h->upload(fd_coeffs(h->r));
        d.fT = h->dalloc<double>((size_t)nm * 3 * N);
        d.fX = h->dalloc<double>((size_t)nm * 3 * N);
        d.fP = h->dalloc<cd>((size_t)nm * 5 * N);
        d.fU = h->dalloc<double>((size_t)3 * N);
    }
    if (d.coll) {
        // get_der_mat, collocation branch (chebyshev.f90:296-321)
        std::vector<double> cm((size_t)N * N), cdm((size_t)N * N), cd2m((size_t)N * N);
        for (int k = 0; k < N; k++) {
            double *r0 = &cm[(size_t)k * N], *r1 = &cdm[(size_t)k * N], *r2 = &cd2m[(size_t)k * N];
            const double xk = h->x[k], a = h->drx[k], b = 0.0;
            r0[0] = 1.0; r0[1] = xk;
            r1[0] = 0.0; r1[1] = a;
            r2[0] = 0.0; r2[1] = b;
            for (int n = 2; n < N; n++) {
                r0[n] = 2.0 * xk * r0[n - 1] - r0[n - 2];
                r1[n] = 2.0 * a * r0[n - 1] + 2.0 * xk * r1[n - 1] - r1[n - 2];
                r2[n] = 2.0 * b * r0[n - 1] + 4.0 * a * r1[n - 1] + 2.0 * xk * r2[n - 1] - r2[n - 2];
            }
        }
        d.cM = h->upload(cm);
        d.cdM = h->upload(cdm);
        d.cd2M = h->upload(cd2m);
        if (!c.l_direct_solve) {
            // influence-matrix method (update_psi_coll_dmat.f90): the
            // inverses and homogeneous solutions per m, and the wall rows of
            // the velocity condition, rnorm and boundary_fac folded in
            // (get_bc_influence_matrix, :1302-1376)
            d.dmat = 1;
            d.OmI = h->dalloc<double>((size_t)nm * N * d.ldi);
            d.PsI = h->dalloc<double>((size_t)nm * N * d.ldi);
            d.dhom = h->dalloc<double>((size_t)nm * 4 * N);
            d.dinfl = h->dalloc<double>((size_t)nm * 4);
            std::vector<double> bc(2 * (size_t)N);
            for (int n = 0; n < N; n++) {
                const double fac = (n == 0 || n == N - 1) ? 0.5 : 1.0, rn = std::sqrt(2.0 / (double)(N - 1));
                const size_t t = n, b = (size_t)(N - 1) * N + n;
                bc[n] = rn * fac * (c.ktopv == 1 ? cd2m[t] - h->or1[0] * cdm[t] : cdm[t]);
                bc[N + n] = rn * fac * (c.kbotv == 1 ? cd2m[b] - h->or1[N - 1] * cdm[b] : cdm[b]);
            }
            d.bcrow = h->upload(bc);
        }
        d.PC_inv = h->dalloc<cd>((size_t)nm * N * 2 * N);
        d.U_inv = h->dalloc<double>((size_t)N * d.ldi);
        // Gauss-Jordan blocks per m: T and xi (N x 2N real), psi (2N x 4N complex)
        d.coll_ws = (2 * (size_t)N * 2 * N * sizeof(double) + (size_t)2 * N * 4 * N * sizeof(cd) + 255) & ~(size_t)255;
        d.coll_work = h->dalloc<char>((size_t)nm * d.coll_ws);
        if (h->sch.l_assembly) {
            // boundary first-derivative rows for robin_bc (chebyshev.f90:338-360):
            // the flipped-row form at r_cmb, its mirror at r_icb, times the map
            std::vector<double> top(N), bot(N);
            top[0] = (2.0 * (N - 1) * (N - 1) + 1.0) / 6.0;
            for (int k = 2; k <= N; k++) {
                const double sn = std::sin(0.5 * (k - 1) * M_PI / (N - 1));
                top[k - 1] = 2.0 * ((k % 2 == 0) ? -1.0 : 1.0) / (2.0 * sn * sn);
            }
            top[N - 1] *= 0.5;
            for (int k = 0; k < N; k++) bot[k] = -top[N - 1 - k];
            const double dt0 = h->drx[0], db0 = h->drx[N - 1];
            for (int k = 0; k < N; k++) { top[k] *= dt0; bot[k] *= db0; }
            d.dr_top = h->upload(top);
            d.dr_bot = h->upload(bot);
            d.A_inv = h->dalloc<double>((size_t)nm * N * d.ldi);
        }
    }
    if (!d.coll && !d.fd) build_operators(h);
    // LDS budgets
    h->lds_radial = (5 * (size_t)h->nm + 7 * (size_t)h->nphi) * sizeof(cd) + 16 * sizeof(double);
    h->mthreads = mloop_threads(N);
    h->lds_mphase = MpLayout(N, h->mthreads).total;
    h->lds_fd = FdLayout(N).total;
    h->lds_fd_loop = FdLoopLayout(N).total;
    h->lds_start = 15 * (size_t)N * sizeof(cd);
    h->lds_start_coll = (13 * (size_t)N + 6 * (size_t)(2 * N - 2) + 3 * (size_t)N + 16) * sizeof(cd) +
                        3 * (size_t)N * sizeof(double);
    {
        size_t nbp = N - 4;
        size_t psi = (size_t)(PZ_P_LD * nbp + 2 * PZ_P_NT * nbp + 16) * sizeof(cd) + N * sizeof(double) +
                     nbp * sizeof(int) + nbp * sizeof(cd) + 7 * 16;
        h->lds_build = 4 * (size_t)N * sizeof(double) + psi + 64;
    }
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    size_t lim = prop.sharedMemPerBlock;
    // rebuild panels of 16 pivots where they fit next to the larger static LDS
    // of the single-run and batched rebuild kernels, else 8
    const size_t gj_static = gj_static_lds();
    d.gj_b = gj_lds_bytes<cd>(2 * N, 16) + gj_static <= lim ? 16 : 8;
    h->lds_gj = gj_lds_bytes<cd>(2 * N, d.gj_b);
    if (d.coll && h->lds_mphase > lim) {
        h->mcoll_l2 = true;
        h->lds_mphase = MpLayout(N, h->mthreads, false).total;
    }
    if (h->lds_radial > lim) {
        h->radial_big = true;
        h->lds_radial = 4 * (size_t)h->nphi * sizeof(cd) + 16 * sizeof(double);
        // without composition: two columns and in-place transforms where the
        // plan allows (k_radial_loop_inplace), the spectra by radius from
        // k_gather_r
        // (the plan then starts with its largest odd prime, fft_inplace's
        // twiddle-free conjugate-pair stage, and takes 2 x 3 as radix 6)
        int plan[PZ_MAX_RADICES];
        for (int s = 0; s < d.nrad; s++) plan[s] = d.radices[s];
        const int nplan = inplace_plan(plan, d.nrad);
        if (!d.l_chem && fft_inplace_fits(h->nphi, 2, plan, nplan, PZ_RADIAL_INPLACE_THREADS)) {
            d.nrad = nplan;
            for (int s = 0; s < d.nrad; s++) d.radices[s] = plan[s];
            h->radial_inplace = true;
            h->lds_radial = (2 * (size_t)h->nphi + 64 + (h->nphi + 63) / 64 + PZ_FFT_OM_MAX) * sizeof(cd) +
                            32 * sizeof(double);
            d.rrow = h->dalloc<cd>(4 * (size_t)h->nm * N);
        }
    }
    // waves of k_build_inv: as many right-hand sides as fit next to the factors
    while (h->inv_waves > 1 && InvLayout(N, h->inv_waves).total > lim) h->inv_waves--;
    h->lds_inv = InvLayout(N, h->inv_waves).total;
    // the kernels of this run's radial method
    size_t need = h->lds_radial;
    if (d.fd) need = std::max({need, h->lds_fd, h->lds_fd_loop});
    else if (d.coll) need = std::max({need, h->lds_mphase, h->lds_start_coll, h->lds_gj + gj_static,
                                      h->sch.l_assembly ? asm_lds_bytes(N) : (size_t)0});
    else need = std::max({need, h->lds_mphase, h->lds_build, h->lds_start, h->lds_inv});
    if (need > lim) throw std::invalid_argument("problem size exceeds the LDS budget of one workgroup");
    allow_full_lds((int)lim);
    HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    h->scratch = scratch_address();
    // device copy of the (now final) descriptor for kernels that read it
    // through a pointer
    h->ddev = h->dalloc<PzDev>(1);
    HIPCHK(hipMemcpy(h->ddev, &h->d, sizeof(PzDev), hipMemcpyHostToDevice));
}

extern "C" int pz_create(const pz_config *cfg, pz_handle **out)
{
    if (!cfg || !out) return -1;
    *out = nullptr;
    pz_handle *h = new pz_handle();
    h->c = *cfg;
    const pz_config &c = *cfg;
    std::string why;
    if (!c.l_direct_solve && (!c.l_cheb_coll || c.l_finite_diff))
        why = "matrix_solve='influence' is implemented for the collocation method";
    else if (c.l_non_rot && !c.l_cheb_coll && !c.l_finite_diff)
        why = "non-rotating runs are implemented for the collocation and finite-difference methods";
    else if (c.radratio <= 0.0) why = "full-disk geometry is not implemented";
    else if (c.n_r_max < 8 || c.n_r_max > 257) why = "n_r_max must be in [8, 257]";
    else if (c.l_cheb_coll && !c.l_finite_diff && c.n_r_max > 256) why = "the collocation method needs n_r_max <= 256";
    else if (c.n_cheb_max < 1 || c.n_cheb_max > c.n_r_max) why = "n_cheb_max must be in [1, n_r_max]";
    else if (c.m_max < 3 || c.minc < 1 || c.m_max % c.minc) why = "invalid m_max/minc";
    else {
        Scheme probe;
        if (!make_scheme(c.time_scheme, c.courfac, probe)) why = "time scheme not implemented";
        else if (probe.l_assembly && !c.l_cheb_coll && !c.l_finite_diff)
            why = "DIRK schemes with an assembly stage are implemented for the collocation and finite-difference "
                  "methods only";
        else if (probe.l_assembly && !c.l_direct_solve)
            why = "DIRK schemes with an assembly stage are implemented for matrix_solve='direct' only";
    }
    if (!why.empty()) {
        int rc = fail(h, -2, why);
        std::fprintf(stderr, "pz_create: %s\n", why.c_str());
        delete h;
        return rc;
    }
    try {
        setup(h);
    } catch (std::exception &e) {
        std::fprintf(stderr, "pz_create: %s\n", e.what());
        delete h;
        return -3;
    }
    *out = h;
    return 0;
}

extern "C" int pz_destroy(pz_handle *h)
{
    delete h;
    return 0;
}

extern "C" int pz_dims(const pz_handle *h, int *nm, int *nr, int *np)
{
    if (!h) return -1;
    *nm = h->nm;
    *nr = h->N;
    *np = h->nphi;
    return 0;
}

extern "C" void *pz_stream(pz_handle *h) { return h ? (void *)h->stream : nullptr; }

static cd *field_ptr(pz_handle *h, int f)
{
    PzDev &d = h->d;
    switch (f) {
    case PZ_US: return d.us;
    case PZ_UP: return d.up;
    case PZ_OM: return d.om;
    case PZ_PSI: return d.psi;
    case PZ_TEMP: return d.temp;
    case PZ_XI: return d.xi;
    case PZ_TEMP_HAT: return d.temp_hat;
    case PZ_XI_HAT: return d.xi_hat;
    case PZ_PSI_HAT: return d.psi_hat;
    case PZ_DTEMP: return d.dtemp;
    default: return nullptr;
    }
}

// Fortran layout (n_m, n_r), m fastest  <->  device [m][r]
static void to_device_layout(const pz_handle *h, const double *host, std::vector<double2> &v)
{
    v.resize((size_t)h->nm * h->N);
    for (int k = 0; k < h->N; k++)
        for (int i = 0; i < h->nm; i++) {
            size_t s = 2 * ((size_t)k * h->nm + i);
            v[(size_t)i * h->N + k] = make_double2(host[s], host[s + 1]);
        }
}
static void from_device_layout(const pz_handle *h, const std::vector<double2> &v, double *host)
{
    for (int k = 0; k < h->N; k++)
        for (int i = 0; i < h->nm; i++) {
            size_t s = 2 * ((size_t)k * h->nm + i);
            host[s] = v[(size_t)i * h->N + k].x;
            host[s + 1] = v[(size_t)i * h->N + k].y;
        }
}

extern "C" int pz_set_field(pz_handle *h, int f, const double *host)
{
    if (!h || !host) return -1;
    cd *p = field_ptr(h, f);
    if (!p) return fail(h, -1, "unknown field");
    try {
        std::vector<double2> v;
        to_device_layout(h, host, v);
        HIPCHK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(cd), hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    } catch (std::exception &e) {
        return fail(h, -4, e.what());
    }
    return 0;
}

extern "C" int pz_get_field(pz_handle *h, int f, double *host)
{
    if (!h || !host) return -1;
    cd *p = field_ptr(h, f);
    if (!p) return fail(h, -1, "unknown field");
    try {
        std::vector<double2> v((size_t)h->nm * h->N);
        HIPCHK(hipMemcpyAsync(v.data(), p, v.size() * sizeof(cd), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        from_device_layout(h, v, host);
    } catch (std::exception &e) {
        return fail(h, -4, e.what());
    }
    return 0;
}

// Gaspari-Cohn pieces (useful.f90:295-341)
// ---------------------------------------------------------------------------
// Checkpoint and frame files (checkpoints.f90, output_frames.f90)
// ---------------------------------------------------------------------------
static std::vector<double2> dev_to_host(pz_handle *h, const cd *p, size_t n)
{
    std::vector<double2> v(n);
    HIPCHK(hipMemcpyAsync(v.data(), p, n * sizeof(cd), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return v;
}

// device [m][r] -> stored f(n_m, n_r), m fastest
static void put_field(pzck::Writer &w, const pz_handle *h, const std::vector<double2> &v)
{
    std::vector<double2> o(v.size());
    for (int i = 0; i < h->nm; i++)
        for (int k = 0; k < h->N; k++) o[(size_t)k * h->nm + i] = v[(size_t)i * h->N + k];
    w.raw(o.data(), o.size() * sizeof(double2));
}

extern "C" int pz_write_checkpoint(pz_handle *h, const char *path)
{
    if (!h || !path) return -1;
    if (!h->started) return fail(h, -5, "pz_start has not been called");
    try {
        PzScheme s;
        HIPCHK(hipMemcpyAsync(&s, h->d.ss, sizeof(s), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        const Scheme &sc = h->sch;
        const size_t F = (size_t)h->nm * h->N;
        const bool fd = h->d.fd;
        pzck::Writer w(path);
        // header (checkpoints.f90:99-134)
        w.put<int>(7);
        w.put<double>(s.time);
        w.put<int>(h->c.l_cheb_coll ? 1 : 0);
        w.text(sc.family == 0 ? "MULTISTEP" : "DIRK", 10);
        w.put<int>(sc.nexp);
        w.put<int>(sc.nimp);
        w.put<int>(sc.nold);
        for (int i = 0; i < (sc.family == 0 ? sc.nexp : 1); i++) w.put<double>(s.dt[i]);
        for (double v : {h->c.ra, h->c.pr, h->c.raxi, h->c.sc, h->d.ek, 0.0 /* stef */, h->c.radratio}) w.put<double>(v);
        w.put<int>(h->N);
        w.put<int>(h->mmax);
        w.put<int>(h->minc);
        w.text(fd ? "fd" : "cheb", 72);
        w.put<int>(fd ? h->N : h->ncheb); // rscheme%n_max
        w.put<int>(fd ? 2 : 0);           // rscheme%order_boundary (no mapping)
        w.put<double>(0.8);               // alph1, alph2: namelist defaults (namelists.f90:556-557)
        w.put<double>(0.0);
        w.raw(h->r.data(), sizeof(double) * h->N);
        w.put<int>(h->l_heat ? 1 : 0);
        w.put<int>(h->l_chem ? 1 : 0);
        w.put<int>(0); // no phase field
        // fields and their history (write_one_field_mpi, checkpoints.f90:868-924)
        auto history = [&](cd *old, cd *impl, cd *expl) {
            if (sc.family != 0) return;
            for (int o = 1; o < sc.nexp; o++) put_field(w, h, dev_to_host(h, expl + o * F, F));
            for (int o = 1; o < sc.nimp; o++) put_field(w, h, dev_to_host(h, impl + o * F, F));
            for (int o = 1; o < sc.nold; o++) put_field(w, h, dev_to_host(h, old + o * F, F));
        };
        put_field(w, h, dev_to_host(h, h->d.us, F));
        put_field(w, h, dev_to_host(h, h->d.up, F));
        history(h->d.P_old, h->d.P_impl, h->d.P_expl);
        if (h->l_heat) {
            put_field(w, h, dev_to_host(h, h->d.temp, F));
            history(h->d.T_old, h->d.T_impl, h->d.T_expl);
        }
        if (h->l_chem) {
            put_field(w, h, dev_to_host(h, h->d.xi, F));
            history(h->d.X_old, h->d.X_impl, h->d.X_expl);
        }
        w.close();
    } catch (std::exception &e) {
        return fail(h, -4, e.what());
    }
    return 0;
}

extern "C" int pz_read_checkpoint(pz_handle *h, const char *path)
{
    if (!h || !path) return -1;
    try {
        pzck::Reader in(path);
        const pzck::Header hd = pzck::read_header(in);
        const Scheme &sc = h->sch;
        const size_t F = (size_t)h->nm * h->N, Fo = (size_t)hd.n_m() * hd.n_r;
        // the dt array of this run's scheme (checkpoints.f90:278-282, 357-366)
        const int nnew = sc.family == 0 ? sc.nexp : 1;
        std::vector<double> dto;
        if (hd.dirk()) dto.assign(std::max(1, nnew), hd.dt[0]);
        else {
            dto.assign(std::max((int)hd.dt.size(), nnew), 0.0);
            for (size_t i = 0; i < hd.dt.size(); i++) dto[i] = hd.dt[i];
            for (int o = (int)hd.dt.size() - 1; o < nnew; o++) dto[o] = hd.dt.back();
        }
        for (int o = 0; o < 8; o++) h->rst.dt[o] = o < (int)dto.size() ? dto[o] : dto.back();
        if (sc.family == 0)
            for (int o = nnew; o < 8; o++) h->rst.dt[o] = h->rst.dt[nnew - 1];
        // restart flags (checkpoints.f90:398-434)
        int nimp_old = hd.nimp, nold_old = hd.nold;
        if (hd.version <= 5) { nimp_old--; nold_old--; }
        h->rst.ab1 = hd.ek != h->d.ek;
        if (sc.family == 0) {
            h->rst.bridge = hd.dirk() ? 1 : (sc.nold > nold_old || sc.nimp > nimp_old);
            if (hd.l_coll != (h->c.l_cheb_coll != 0)) h->rst.ab1 = h->rst.bridge = 1;
            if (hd.version < 4 && nold_old > 2) h->rst.ab1 = h->rst.bridge = 1;
        } else {
            h->rst.bridge = 0;
        }
        pzck::Target t;
        t.N = h->N;
        t.nm = h->nm;
        t.idx2m = h->idx2m.data();
        t.r = h->r.data();
        t.cheb = !h->d.fd;
        t.order_boundary = h->d.fd ? 2 : 0;
        t.rnorm = h->d.fd ? 1.0 : h->rnorm;
        std::vector<double2> raw(Fo), mapped;
        auto next = [&](cd *dst, bool lBc, bool phys) {
            in.raw(raw.data(), Fo * sizeof(double2));
            if (!dst) return;
            pzck::map_field(hd, raw, t, lBc, phys, mapped);
            HIPCHK(hipMemcpyAsync(dst, mapped.data(), F * sizeof(cd), hipMemcpyHostToDevice, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
        };
        auto history = [&](cd *old, cd *impl, cd *expl) {
            if (hd.dirk()) return;
            const bool ms = sc.family == 0;
            for (int o = 2; o <= hd.nexp; o++) next(ms && o <= sc.nexp ? expl + (o - 1) * F : nullptr, true, hd.l_coll);
            for (int o = 2; o <= nimp_old; o++) next(ms && o <= sc.nimp ? impl + (o - 1) * F : nullptr, true, hd.l_coll);
            if (hd.version > 3)
                for (int o = 2; o <= nold_old; o++) next(ms && o <= sc.nold ? old + (o - 1) * F : nullptr, false, hd.l_coll);
        };
        next(h->d.us, false, true);
        next(h->d.up, false, true);
        history(h->d.P_old, h->d.P_impl, h->d.P_expl);
        if (hd.l_heat) {
            next(h->d.temp, false, true);
            history(h->d.T_old, h->d.T_impl, h->d.T_expl);
        }
        if (hd.l_chem) {
            next(h->d.xi, false, true);
            history(h->d.X_old, h->d.X_impl, h->d.X_expl);
        }
        if (!h->d.coll && !h->d.fd) {
            // the integration method rebuilds psi(m /= 0) from the stored u_s
            // (init_fields.f90:67-95); the finite-difference start does the
            // same in its start kernel, collocation does not need psi
            std::vector<double2> us = dev_to_host(h, h->d.us, F), psi(F, double2{0.0, 0.0});
            const double rc2 = h->r_cmb * h->r_cmb;
            for (int i = 0; i < h->nm; i++) {
                const int m = h->idx2m[i];
                if (m == 0) continue;
                for (int k = 1; k < h->N; k++) {
                    const double f = h->r[k] / (double)m / (rc2 - h->r[k] * h->r[k]);
                    const double2 u = us[(size_t)i * h->N + k];
                    psi[(size_t)i * h->N + k] = double2{f * u.y, -f * u.x}; // -i f u
                }
            }
            HIPCHK(hipMemcpyAsync(h->d.psi, psi.data(), F * sizeof(cd), hipMemcpyHostToDevice, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
        }
        h->rst.time = hd.time;
        h->rst.pending = true;
        h->started = false;
    } catch (std::exception &e) {
        return fail(h, -4, e.what());
    }
    return 0;
}

extern "C" int pz_read_checkpoint_header(const char *path, pz_checkpoint_info *info, double *r, int r_len)
{
    if (!path || !info) return -1;
    try {
        pzck::Reader in(path);
        const pzck::Header hd = pzck::read_header(in);
        std::memset(info, 0, sizeof(*info));
        info->version = hd.version;
        info->time = hd.time;
        info->l_cheb_coll = hd.l_coll;
        std::snprintf(info->family, sizeof(info->family), "%s", hd.family.substr(0, hd.family.find(' ')).c_str());
        info->nexp = hd.nexp;
        info->nimp = hd.nimp;
        info->nold = hd.nold;
        info->n_dt = (int)std::min<size_t>(hd.dt.size(), 16);
        for (int i = 0; i < info->n_dt; i++) info->dt[i] = hd.dt[i];
        info->ra = hd.ra;
        info->pr = hd.pr;
        info->raxi = hd.raxi;
        info->sc = hd.sc;
        info->ek = hd.ek;
        info->stef = hd.stef;
        info->radratio = hd.radratio;
        info->n_r_max = hd.n_r;
        info->m_max = hd.m_max;
        info->minc = hd.minc;
        std::snprintf(info->rscheme, sizeof(info->rscheme), "%s", hd.rscheme.substr(0, hd.rscheme.find(' ')).c_str());
        info->n_max = hd.n_max;
        info->order_boundary = hd.order_boundary;
        info->alph1 = hd.alph1;
        info->alph2 = hd.alph2;
        info->l_heat = hd.l_heat;
        info->l_chem = hd.l_chem;
        info->l_phase_field = hd.l_phase;
        info->header_bytes = hd.header_bytes;
        info->n_arrays = hd.n_arrays();
        if (r) for (int i = 0; i < std::min(r_len, hd.n_r); i++) r[i] = hd.r[i];
    } catch (std::exception &) {
        return -4;
    }
    return 0;
}

// write_snapshot_mloc (output_frames.f90:96-167), version 2
extern "C" int pz_write_frame(pz_handle *h, int field, const char *path)
{
    if (!h || !path) return -1;
    cd *p = field_ptr(h, field);
    if (!p) return fail(h, -1, "unknown field");
    try {
        PzScheme s;
        HIPCHK(hipMemcpyAsync(&s, h->d.ss, sizeof(s), hipMemcpyDeviceToHost, h->stream));
        std::vector<double2> v = dev_to_host(h, p, (size_t)h->nm * h->N);
        std::vector<double> tc(h->N), xc(h->N);
        HIPCHK(hipMemcpy(tc.data(), h->d.tcond, sizeof(double) * h->N, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(xc.data(), h->d.xicond, sizeof(double) * h->N, hipMemcpyDeviceToHost));
        pzck::Writer w(path);
        w.put<int>(2);
        for (double x : {h->started ? s.time : 0.0, h->c.ra, h->d.ek, h->c.pr, h->c.radratio, h->c.sc, h->c.raxi})
            w.put<double>(x);
        for (int x : {h->N, h->nm, h->mmax, h->minc, h->nphi}) w.put<int>(x);
        w.raw(h->r.data(), sizeof(double) * h->N);
        w.raw(tc.data(), sizeof(double) * h->N);
        w.raw(xc.data(), sizeof(double) * h->N);
        w.raw(h->idx2m.data(), sizeof(int) * h->nm);
        put_field(w, h, v);
        w.close();
    } catch (std::exception &e) {
        return fail(h, -4, e.what());
    }
    return 0;
}

static double gc_center(double c, double L)
{
    const double pi = 4.0 * std::atan(1.0);
    return pi * L * L * L * (1.0 - std::exp(-2.0 * c / L)) - 2.0 * pi * c * L * (c + L) * std::exp(-2.0 * c / L);
}
static double gc_middle(double r, double c, double L)
{
    const double pi = 4.0 * std::atan(1.0);
    return pi / 3.0 * L * r * (r + 3.0 * L) * std::exp(-r / L) +
           (2.0 * pi * L * L * (c + L) * (c + L)) / r * std::exp(-2.0 * c / L) * (1.0 - (1.0 - r / (c + L)) * std::exp(r / L)) +
           pi * L * L * L * (std::exp(-r / L) - std::exp((r - 2.0 * c) / L));
}
static double gc_edge(double r, double c, double L)
{
    const double pi = 4.0 * std::atan(1.0);
    return 2.0 * pi * L / r * std::exp(-r / L) *
           (0.5 * (r * (r + L) * (2.0 * c - r)) + 1.0 / 3.0 * (std::pow(r - c, 3) - c * c * c) - L * (c + L) * (r - c + L) +
            L * (c + L) * (c + L) * std::exp((r - 2.0 * c) / L));
}

// initProp (init_fields.f90:260-360), deterministic branches; adds to v[m][r]
static int init_prop(pz_handle *h, std::vector<double2> &v, int init, double amp)
{
    const int N = h->N;
    const double pi = 4.0 * std::atan(1.0);
    if (init > 0) {
        if (init % h->minc != 0 || init > h->mmax) return -3;
        int idx = init / h->minc;
        for (int n = 0; n < N; n++) v[(size_t)idx * N + n].x += std::sin(pi * (h->r[n] - h->r_icb)) * amp;
        return 0;
    }
    if (init != -1) return -4; // the unseeded random branch is not reproducible
    const double sigma_r = 0.1 / std::sqrt(2.0);
    double rcm = 0.5 * (h->r_cmb + h->r_icb);
    int ir = 0;
    for (int n = 1; n < N; n++)
        if (std::fabs(h->r[n] - rcm) < std::fabs(h->r[ir] - rcm)) ir = n;
    rcm = h->r[ir];
    const double c1 = 0.5 * (h->r[0] - rcm), c2 = 0.5 * (rcm - h->r[N - 1]), L = 0.5 * sigma_r;
    std::vector<double> gasp(N, 0.0);
    for (int n = 0; n < N; n++) {
        double r = h->r[n];
        if (n == ir) gasp[n] = 1.0;
        else if (r < rcm - c2) gasp[n] = gc_edge(rcm - r, c2, L);
        else if (r >= rcm - c2 && r < rcm) gasp[n] = gc_middle(rcm - r, c2, L);
        else if (r > rcm && r <= rcm + c1) gasp[n] = gc_middle(r - rcm, c1, L);
        else if (r > rcm + c1) gasp[n] = gc_edge(r - rcm, c1, L);
    }
    for (int n = 0; n < N; n++) {
        if (n < ir) gasp[n] /= gc_center(c1, L);
        else if (n > ir) gasp[n] /= gc_center(c2, L);
    }
    // fourier::fft of the azimuthal profile (1/n_phi normalisation)
    const int np = h->nphi;
    const double phi0 = pi / h->minc;
    std::vector<double> prof(np);
    for (int j = 0; j < np; j++) {
        double phi = (double)j * 2.0 * pi / h->minc / (double)np;
        prof[j] = std::exp(-(phi - phi0) * (phi - phi0) / (0.2 * 0.2));
    }
    const ld lpi = 3.14159265358979323846264338327950288L;
    std::vector<double2> spec(h->nm);
    for (int i = 0; i < h->nm; i++) {
        ld sr = 0, si = 0;
        for (int j = 0; j < np; j++) {
            ld ang = 2.0L * lpi * (ld)((long)i * j % np) / (ld)np;
            sr += (ld)prof[j] * cosl(ang);
            si -= (ld)prof[j] * sinl(ang);
        }
        spec[i] = make_double2((double)(sr / np), (double)(si / np));
    }
    for (int n = 0; n < N; n++) {
        double cr = amp * gasp[n];
        for (int i = 0; i < h->nm; i++) {
            v[(size_t)i * N + n].x += cr * spec[i].x;
            v[(size_t)i * N + n].y += cr * spec[i].y;
        }
    }
    return 0;
}

extern "C" int pz_init_fields(pz_handle *h, int init_t, double amp_t, int init_xi, double amp_xi, int init_u,
                              double amp_u)
{
    if (!h) return -1;
    try {
        const size_t F = (size_t)h->nm * h->N;
        auto edit = [&](cd *dptr, auto fn) {
            std::vector<double2> v(F);
            HIPCHK(hipMemcpy(v.data(), dptr, F * sizeof(cd), hipMemcpyDeviceToHost));
            int rc = fn(v);
            HIPCHK(hipMemcpy(dptr, v.data(), F * sizeof(cd), hipMemcpyHostToDevice));
            return rc;
        };
        int rc = 0;
        if (init_t != 0 && h->l_heat) rc |= edit(h->d.temp, [&](std::vector<double2> &v) { return init_prop(h, v, init_t, amp_t); });
        if (init_xi != 0 && h->l_chem) rc |= edit(h->d.xi, [&](std::vector<double2> &v) { return init_prop(h, v, init_xi, amp_xi); });
        if (init_u != 0) {
            // initU (init_fields.f90:455-509)
            const double pi = 4.0 * std::atan(1.0);
            const int N = h->N;
            if (init_u > 0) {
                if (init_u % h->minc != 0 || init_u > h->mmax) return fail(h, -3, "init_u out of range");
                int idx = init_u / h->minc;
                rc |= edit(h->d.us, [&](std::vector<double2> &v) {
                    for (int n = 0; n < N; n++) v[(size_t)idx * N + n].x += std::sin(pi * (h->r[n] - h->r_icb)) * amp_u;
                    return 0;
                });
            } else {
                rc |= edit(h->d.up, [&](std::vector<double2> &v) {
                    for (int n = 0; n < N; n++) v[n].x += amp_u * std::sin(pi * (h->r[n] - h->r_icb));
                    return 0;
                });
            }
        }
        if (rc) return fail(h, rc, "unsupported initial condition");
    } catch (std::exception &e) {
        return fail(h, -4, e.what());
    }
    return 0;
}

extern "C" int pz_start(pz_handle *h)
{
    if (!h) return -1;
    try {
        PzScheme s;
        std::memset(&s, 0, sizeof(s));
        const bool rst = h->rst.pending;
        // a fresh start: dt(:) = dtMax (init_fields.f90:111); a restart: the
        // time and dt array of the checkpoint (checkpoints.f90:357-366)
        for (int i = 0; i < 8; i++) s.dt[i] = rst ? h->rst.dt[i] : h->c.dtMax;
        s.time = rst ? h->rst.time : 0.0;
        s.ab1 = rst ? h->rst.ab1 : 0;
        s.lMatNext = 1; // step_time.f90:115
        HIPCHK(hipMemcpyAsync(h->d.ss, &s, sizeof(s), hipMemcpyHostToDevice, h->stream));
        // l_bridge_step: always on a fresh start, as read_checkpoint decided on
        // a restart (the kernels read it through the descriptor)
        const int bridge = rst ? (h->bridge_default ? h->rst.bridge : 0) : h->bridge_default;
        if (bridge != h->d.bridge) {
            h->d.bridge = bridge;
            HIPCHK(hipMemcpyAsync(h->ddev, &h->d, sizeof(PzDev), hipMemcpyHostToDevice, h->stream));
        }
        const size_t F = (size_t)h->nm * h->N;
        const Scheme &sc = h->sch;
        cd *slots[] = {h->d.T_old, h->d.T_impl, h->d.T_expl, h->d.X_old, h->d.X_impl, h->d.X_expl,
                       h->d.P_old, h->d.P_impl, h->d.P_expl};
        size_t cnt[] = {(size_t)sc.nold, (size_t)sc.nimp, (size_t)sc.nexp, (size_t)sc.nold, (size_t)sc.nimp,
                        (size_t)sc.nexp, (size_t)sc.nold, (size_t)sc.nimp, (size_t)sc.nexp};
        // a restart keeps the history slots 2.. that the checkpoint filled
        for (int i = 0; i < 9; i++)
            HIPCHK(hipMemsetAsync(slots[i], 0, (rst ? 1 : cnt[i]) * F * sizeof(cd), h->stream));
        hipLaunchKernelGGL(k_set_weights, dim3(1), dim3(1), 0, h->stream, h->d);
        // exactly one start kernel per radial method: k_start reads the banded
        // operators that only the integration method builds
        if (h->d.fd) {
            hipLaunchKernelGGL(k_fd_start, dim3(h->nm), dim3(256), h->lds_fd, h->stream, (const PzDev *)h->ddev);
        } else if (h->d.coll) {
            hipLaunchKernelGGL(k_start_coll, dim3(h->nm), dim3(256), h->lds_start_coll, h->stream,
                               (const PzDev *)h->ddev);
            if (h->d.l_assembly) {
                if (!h->d.A_inv || !h->d.dr_top || !h->d.dr_bot)
                    throw std::logic_error("assembly stage without its matrices");
                hipLaunchKernelGGL(k_build_ass, dim3(h->nm), dim3(1024), 0, h->stream, h->d);
            }
        } else {
            if (!h->d.T_RHSE2 || !h->d.T_RHSI || !h->d.P_RHSI)
                throw std::logic_error("integration start without its banded operators");
            hipLaunchKernelGGL(k_start, dim3(h->nm), dim3(256), h->lds_start, h->stream, h->d);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(h->stream));
        if (h->gexec) { (void)hipGraphExecDestroy(h->gexec); h->gexec = nullptr; }
        if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
        h->started = true;
        h->restarted = rst;
        h->rst.pending = false;
    } catch (std::exception &e) {
        return fail(h, -4, e.what());
    }
    return 0;
}

// The pieces of one stage of time_loop (step_time.f90:215-370).
// Explicit part: transposes + radial_loop (step_time.f90:222-251), when the
// scheme evaluates the explicit terms at this stage.
static void launch_radial(pz_handle *h, int is, int courant)
{
    if (h->radial_inplace) {
        hipLaunchKernelGGL(k_gather_r, dim3((h->nm + PZ_GATHER_TM - 1) / PZ_GATHER_TM, (h->N + PZ_GATHER_TR - 1) / PZ_GATHER_TR),
                           dim3(256), 0, h->stream, (const PzDev *)h->ddev);
        hipLaunchKernelGGL(k_radial_loop_inplace, dim3(h->N), dim3(PZ_RADIAL_INPLACE_THREADS), h->lds_radial, h->stream,
                           (const PzDev *)h->ddev, is, courant);
    } else {
        hipLaunchKernelGGL(h->radial_big ? k_radial_loop_big : k_radial_loop, dim3(h->N),
                           dim3(h->radial_big ? PZ_RADIAL_BIG_THREADS : 256), h->lds_radial, h->stream,
                           (const PzDev *)h->ddev, is, courant);
    }
}
static void launch_explicit(pz_handle *h, int stage)
{
    if (h->sch.l_exp_calc[stage - 1]) launch_radial(h, stage - 1, stage == 1 ? 1 : 0);
}

// Stage 1: dt_courant, the dt roll, set_weights and the rebuild flag
// (step_time.f90:256-304, courant.f90:78-167)
static void launch_courant(pz_handle *h)
{
    hipLaunchKernelGGL(k_courant, dim3(1), dim3(256), 0, h->stream, (const PzDev *)h->ddev);
}

__global__ void k_force_lmat(PzScheme *s)
{
    if (threadIdx.x == 0) s->lMat = 1;
}

// Stage 1: the matrix rebuild of the run's radial method; the kernels exit
// at once unless the rebuild flag is set
static void launch_builds(pz_handle *h)
{
    if (h->d.fd) {
        hipLaunchKernelGGL(k_fd_build, dim3(h->nm), dim3(256), h->lds_fd, h->stream, (const PzDev *)h->ddev, 0);
    } else if (h->d.coll) {
        hipLaunchKernelGGL(k_build_coll, dim3(h->nm, h->d.dmat ? 4 : 3), dim3(1024), h->lds_gj, h->stream,
                           (const PzDev *)h->ddev, 0);
        if (h->d.dmat)
            hipLaunchKernelGGL(k_dmat_homog, dim3(h->nm), dim3(256), dmat_homog_lds(h->N, h->d.L), h->stream,
                               (const PzDev *)h->ddev, 0);
    } else {
        hipLaunchKernelGGL(k_build_lhs, dim3(3, h->nm), dim3(64), h->lds_build, h->stream, (const PzDev *)h->ddev, 0);
        hipLaunchKernelGGL(k_build_inv, dim3(h->nm, 3), dim3(64 * h->inv_waves), h->lds_inv, h->stream,
                           (const PzDev *)h->ddev, 0);
    }
}

// Implicit part: finish_explicit_assembly (fused into the per-m kernel,
// mloop.f90:169-229) and mloop (mloop.f90:36-103) of this stage; on the
// last stage of a scheme with an assembly stage, assemble_stage instead
// (step_time.f90:313, 353-368)
static void launch_implicit(pz_handle *h, int stage)
{
    const Scheme &s = h->sch;
    const int exp_calc = s.l_exp_calc[stage - 1];
    if (s.l_assembly && stage == s.nstages) {
        if (h->d.fd)
            hipLaunchKernelGGL(k_assemble_fd, dim3(h->nm), dim3(256), h->lds_fd, h->stream, (const PzDev *)h->ddev);
        else
            hipLaunchKernelGGL(k_assemble_coll, dim3(h->nm), dim3(ASM_THREADS), asm_lds_bytes(h->N), h->stream,
                               (const PzDev *)h->ddev);
        return;
    }
    if (h->d.fd)
        hipLaunchKernelGGL(k_fd_mloop, dim3(h->nm), dim3(256), h->lds_fd_loop, h->stream, (const PzDev *)h->ddev, stage,
                           exp_calc);
    else if (h->d.coll)
        hipLaunchKernelGGL(h->mcoll_l2 ? k_mcoll_l2 : k_mcoll, dim3(h->nm), dim3(h->mthreads), h->lds_mphase, h->stream,
                           (const PzDev *)h->ddev, stage, exp_calc);
    else
        hipLaunchKernelGGL(k_mphase, dim3(h->nm), dim3(h->mthreads), h->lds_mphase, h->stream, (const PzDev *)h->ddev,
                           stage, exp_calc);
}

// the stage loop of time_loop (step_time.f90:215-370) for one step
static void launch_step(pz_handle *h)
{
    for (int stage = 1; stage <= h->sch.nstages; stage++) {
        launch_explicit(h, stage);
        if (stage == 1) {
            launch_courant(h);
            launch_builds(h);
        }
        launch_implicit(h, stage);
    }
}

// nsteps steps through the captured step graph (captured on first use)
static void run_steps(pz_handle *h, int nsteps)
{
    if (h->use_graph && !h->gexec) {
        HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        launch_step(h);
        HIPCHK(hipStreamEndCapture(h->stream, &h->graph));
        HIPCHK(hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
    }
    for (int i = 0; i < nsteps; i++) {
        if (h->gexec) HIPCHK(hipGraphLaunch(h->gexec, h->stream));
        else launch_step(h);
    }
    HIPCHK(hipGetLastError());
}

extern "C" int pz_step(pz_handle *h, int nsteps)
{
    if (!h) return -1;
    if (!h->started) return fail(h, -5, "pz_start has not been called");
    try {
        run_steps(h, nsteps);
    } catch (std::exception &e) {
        return fail(h, -4, e.what());
    }
    return 0;
}

// ---- the stage-granular boundary: time_loop's procedures one by one ----
static int stage_check(pz_handle *h, int stage)
{
    if (!h) return -1;
    if (!h->started) return fail(h, -5, "pz_start has not been called");
    if (stage < 1 || stage > h->sch.nstages) return fail(h, -1, "stage out of range");
    return 0;
}

extern "C" int pz_nstages(const pz_handle *h) { return h ? h->sch.nstages : -1; }

extern "C" int pz_stage_explicit(pz_handle *h, int stage)
{
    int rc = stage_check(h, stage);
    if (rc) return rc;
    try {
        launch_explicit(h, stage);
        HIPCHK(hipGetLastError());
    } catch (std::exception &e) {
        return fail(h, -4, e.what());
    }
    return 0;
}

extern "C" int pz_courant(pz_handle *h, double *dtr, double *dth)
{
    int rc = stage_check(h, 1);
    if (rc) return rc;
    try {
        launch_courant(h);
        HIPCHK(hipGetLastError());
        if (dtr || dth) {
            PzScheme s;
            HIPCHK(hipMemcpyAsync(&s, h->d.ss, sizeof(s), hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            if (dtr) *dtr = s.dt_r;
            if (dth) *dth = s.dt_h;
        }
    } catch (std::exception &e) {
        return fail(h, -4, e.what());
    }
    return 0;
}

extern "C" int pz_stage_implicit(pz_handle *h, int stage, int lMat)
{
    int rc = stage_check(h, stage);
    if (rc) return rc;
    try {
        if (stage == 1) {
            if (lMat) hipLaunchKernelGGL(k_force_lmat, dim3(1), dim3(64), 0, h->stream, h->d.ss);
            launch_builds(h);
        }
        launch_implicit(h, stage);
        HIPCHK(hipGetLastError());
    } catch (std::exception &e) {
        return fail(h, -4, e.what());
    }
    return 0;
}

static int check_state(pz_handle *h, PzScheme *out)
{
    PzScheme s;
    HIPCHK(hipMemcpyAsync(&s, h->d.ss, sizeof(s), hipMemcpyDeviceToHost, h->stream));
    std::vector<int> info(4 * h->nm);
    HIPCHK(hipMemcpyAsync(info.data(), h->d.info, info.size() * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (out) *out = s;
    if (s.error) return fail(h, -6, "time step too small (dt < dtMin)");
    for (int v : info)
        if (v) return fail(h, v, "singular matrix in the implicit solve");
    return 0;
}

extern "C" int pz_time(pz_handle *h, double *time, double *dt)
{
    if (!h) return -1;
    try {
        PzScheme s;
        int rc = check_state(h, &s);
        if (time) *time = s.time;
        if (dt) *dt = s.dt[0];
        return rc;
    } catch (std::exception &e) {
        return fail(h, -4, e.what());
    }
}

// rInt_R, finite-difference branch (integration.f90:60-96): Simpson's rule
// on the non-uniform grid, trapezoids at both ends for an even n_r
static double rint_fd(const pz_handle *h, const std::vector<double> &f)
{
    const int N = h->N;
    const std::vector<double> &r = h->r;
    auto simps = [&](int n) {
        const double h2 = r[n + 1] - r[n], h1 = r[n] - r[n - 1];
        return (h1 + h2) / 6.0 *
               (f[n - 1] * (2.0 * h1 - h2) / h1 + f[n] * (h1 + h2) * (h1 + h2) / (h1 * h2) + f[n + 1] * (2.0 * h2 - h1) / h2);
    };
    double s = 0.0;
    if (N % 2 == 1) {
        for (int n = 1; n < N - 1; n += 2) s += simps(n);
        return -s;
    }
    s = 0.5 * (r[1] - r[0]) * (f[1] + f[0]);
    for (int n = 2; n < N - 1; n += 2) s += simps(n);
    s += 0.5 * (r[N - 1] - r[N - 2]) * (f[N - 1] + f[N - 2]);
    for (int n = 1; n < N - 1; n += 2) s += simps(n);
    return -0.5 * s;
}

// rInt_R, Chebyshev branch (integration.f90:18-58)
static double rint_r(const pz_handle *h, const std::vector<double> &f)
{
    if (h->d.fd) return rint_fd(h, f);
    const int N = h->N;
    std::vector<double> f2(N), c(N);
    for (int n = 0; n < N; n++) f2[n] = f[n] / h->drx[n];
    for (int k = 0; k < N; k++) {
        double s = 0.0;
        for (int j = 1; j < N - 1; j++) s += f2[j] * std::cos(4.0 * std::atan(1.0) * (double)((long)j * k % (2 * N - 2)) / (double)(N - 1));
        c[k] = h->cheb_fac * (f2[0] + ((k & 1) ? -1.0 : 1.0) * f2[N - 1] + 2.0 * s);
    }
    c[0] *= 0.5;
    c[N - 1] *= 0.5;
    double s = c[0];
    for (int nc = 3; nc <= N; nc += 2) s -= 1.0 / (double)(nc * (nc - 2)) * c[nc - 1];
    return 2.0 * std::sqrt(2.0 / (double)(N - 1)) * s;
}

extern "C" int pz_ekin_3d(pz_handle *h, double out[4])
{
    if (!h || !out) return -1;
    try {
        const int N = h->N;
        double *dprof = nullptr;
        HIPCHK(hipMallocAsync((void **)&dprof, 3 * N * sizeof(double), h->stream));
        hipLaunchKernelGGL(k_ekin_profiles, dim3((N + 63) / 64), dim3(64), 0, h->stream, h->d, dprof);
        std::vector<double> prof(3 * N);
        HIPCHK(hipMemcpyAsync(prof.data(), dprof, 3 * N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipFreeAsync(dprof, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        // outputs.f90:552-571
        const double pi = 4.0 * std::atan(1.0), tiny = 10.0 * 2.220446049250313e-16;
        auto ro = [&](double v) { return std::fabs(v) < tiny ? 0.0 : v; };
        std::vector<double> t(N);
        for (int n = 0; n < N; n++) t[n] = prof[n] * h->r[n] * h->height[n];
        out[0] = ro(pi * rint_r(h, t));
        for (int n = 0; n < N; n++) t[n] = prof[N + n] * h->r[n] * h->height[n];
        out[1] = ro(pi * rint_r(h, t));
        for (int n = 0; n < N; n++) t[n] = prof[n] * h->beta[n] * h->beta[n] * h->r[n] * std::pow(h->height[n], 3) / 12.0;
        out[2] = ro(pi * rint_r(h, t));
        for (int n = 0; n < N; n++) t[n] = prof[2 * N + n] * h->r[n] * h->height[n];
        out[3] = ro(pi * rint_r(h, t));
    } catch (std::exception &e) {
        return fail(h, -4, e.what());
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Time series and spectra (get_time_series, outputs.f90:418-830;
// get_lengthscales, :897-962; calculate_spectra, spectra.f90:104-197)
// ---------------------------------------------------------------------------
// rInt_R is linear in its integrand: its weights on this run's grid, built
// once from the unit vectors
static const std::vector<double> &rint_weights(pz_handle *h)
{
    if ((int)h->rintw.size() != h->N) {
        std::vector<double> e(h->N, 0.0);
        h->rintw.assign(h->N, 0.0);
        for (int k = 0; k < h->N; k++) {
            e[k] = 1.0;
            h->rintw[k] = rint_r(h, e);
            e[k] = 0.0;
        }
    }
    return h->rintw;
}

static double rint_w(pz_handle *h, const std::vector<double> &f)
{
    const std::vector<double> &w = rint_weights(h);
    double s = 0.0;
    for (int k = 0; k < h->N; k++) s += w[k] * f[k];
    return s;
}

// simps (integration.f90:102-150) on abscissae x
static double simps_x(const std::vector<double> &f, const std::vector<double> &x)
{
    const int n = (int)f.size();
    auto piece = [&](int k) {
        const double h2 = x[k + 1] - x[k], h1 = x[k] - x[k - 1];
        return (h1 + h2) / 6.0 * (f[k - 1] * (2.0 * h1 - h2) / h1 + f[k] * (h1 + h2) * (h1 + h2) / (h1 * h2) +
                                  f[k + 1] * (2.0 * h2 - h1) / h2);
    };
    double s = 0.0;
    if (n % 2 == 1) {
        for (int k = 1; k < n - 1; k += 2) s += piece(k);
        return -s;
    }
    s = 0.5 * (x[1] - x[0]) * (f[1] + f[0]);
    for (int k = 2; k < n - 1; k += 2) s += piece(k);
    s += 0.5 * (x[n - 1] - x[n - 2]) * (f[n - 1] + f[n - 2]);
    for (int k = 1; k < n - 1; k += 2) s += piece(k);
    return -0.5 * s;
}

// the radial derivatives of T and xi as the time series take them: the
// finite-difference ghost-zone stencil of the step, or get_dr of the
// Chebyshev methods
static void series_derivatives(pz_handle *h, cd *dT, cd *dX)
{
    const size_t F = (size_t)h->nm * h->N;
    if (h->d.fd) {
        const int blocks = (int)((F + 255) / 256);
        if (h->l_heat)
            hipLaunchKernelGGL(k_fd_dr_ghost, dim3(blocks), dim3(256), 0, h->stream, h->d, (const cd *)h->d.temp, dT,
                               h->c.ktopt, h->c.kbott, (const cd *)h->d.topt, (const cd *)h->d.bott);
        if (h->l_chem)
            hipLaunchKernelGGL(k_fd_dr_ghost, dim3(blocks), dim3(256), 0, h->stream, h->d, (const cd *)h->d.xi, dX,
                               h->c.ktopxi, h->c.kbotxi, (const cd *)h->d.topxi, (const cd *)h->d.botxi);
    } else {
        const size_t lds = (4 * h->N + 4 * (size_t)h->d.L + 16) * sizeof(cd);
        if (h->l_heat)
            hipLaunchKernelGGL(k_dr_vecs, dim3(h->nm), dim3(256), lds, h->stream, h->d, (const cd *)h->d.temp, dT,
                               (cd *)nullptr, 1);
        if (h->l_chem)
            hipLaunchKernelGGL(k_dr_vecs, dim3(h->nm), dim3(256), lds, h->stream, h->d, (const cd *)h->d.xi, dX,
                               (cd *)nullptr, 1);
    }
}

extern "C" int pz_time_series(pz_handle *h, pz_series *ts)
{
    if (!h || !ts) return -1;
    if (!h->started) return fail(h, -5, "pz_start has not been called");
    try {
        const int N = h->N, nm = h->nm;
        const size_t F = (size_t)nm * N;
        std::memset(ts, 0, sizeof(*ts));
        cd *dd = nullptr;
        double *dprof = nullptr, *dspec = nullptr;
        HIPCHK(hipMallocAsync((void **)&dd, 2 * F * sizeof(cd), h->stream));
        HIPCHK(hipMallocAsync((void **)&dprof, PZ_NPROF * (size_t)N * sizeof(double), h->stream));
        HIPCHK(hipMallocAsync((void **)&dspec, 5 * F * sizeof(double), h->stream));
        HIPCHK(hipMemsetAsync(dd, 0, 2 * F * sizeof(cd), h->stream));
        series_derivatives(h, dd, dd + F);
        hipLaunchKernelGGL(k_series_profiles, dim3((N + 63) / 64), dim3(64), 0, h->stream, h->d, (const cd *)dd,
                           (const cd *)(dd + F), dprof);
        hipLaunchKernelGGL(k_spectra_mr, dim3((int)((F + 255) / 256)), dim3(256), 0, h->stream, h->d, dspec);
        HIPCHK(hipGetLastError());
        std::vector<double> prof(PZ_NPROF * (size_t)N), spec(5 * F);
        std::vector<double2> dT(F), dX(F), T(F), X(F);
        PzScheme s;
        HIPCHK(hipMemcpyAsync(prof.data(), dprof, prof.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(spec.data(), dspec, spec.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(dT.data(), dd, F * sizeof(cd), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(dX.data(), dd + F, F * sizeof(cd), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(T.data(), h->d.temp, F * sizeof(cd), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(X.data(), h->d.xi, F * sizeof(cd), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(&s, h->d.ss, sizeof(s), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipFreeAsync(dd, h->stream));
        HIPCHK(hipFreeAsync(dprof, h->stream));
        HIPCHK(hipFreeAsync(dspec, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));

        const double pi = 4.0 * std::atan(1.0), tiny = 10.0 * 2.220446049250313e-16;
        auto ro = [&](double v) { return std::fabs(v) < tiny ? 0.0 : v; };
        const std::vector<double> &r = h->r, &hg = h->height, &beta = h->beta;
        auto row = [&](int j) { return std::vector<double>(prof.begin() + (size_t)j * N, prof.begin() + (size_t)(j + 1) * N); };
        const std::vector<double> us2 = row(0), up2 = row(1), upa = row(2), enst = row(3), buo = row(4), chem = row(5);
        std::vector<double> nuv = row(6), shv = row(7), flux = row(8);
        std::vector<double> rc(N), t(N);
        // radial functions not kept on the host
        HIPCHK(hipMemcpy(rc.data(), h->d.rgrav, N * sizeof(double), hipMemcpyDeviceToHost));
        std::vector<double> ekp(N), tcond(N), dtcond(N), xicond(N), dxicond(N);
        HIPCHK(hipMemcpy(ekp.data(), h->d.ekpump, N * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(tcond.data(), h->d.tcond, N * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(dtcond.data(), h->d.dtcond, N * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(xicond.data(), h->d.xicond, N * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(dxicond.data(), h->d.dxicond, N * sizeof(double), hipMemcpyDeviceToHost));
        const double CorFac = h->d.CorFac, BuoFac = h->d.BuoFac, ChemFac = h->d.ChemFac, TdiffFac = h->d.TdiffFac;
        const double ViscFac = h->d.ViscFac;
        auto integ = [&](auto fn) {
            for (int k = 0; k < N; k++) t[k] = fn(k);
            return rint_w(h, t);
        };
        ts->time = s.time;
        // kinetic energies (outputs.f90:521-569)
        const double us2_2D = ro(pi * integ([&](int k) { return us2[k] * r[k]; }));
        const double up2_2D = ro(pi * integ([&](int k) { return up2[k] * r[k]; }));
        const double upa_2D = ro(pi * integ([&](int k) { return upa[k] * r[k]; }));
        const double us2_3D = ro(pi * integ([&](int k) { return us2[k] * r[k] * hg[k]; }));
        const double up2_3D = ro(pi * integ([&](int k) { return up2[k] * r[k] * hg[k]; }));
        const double uz2_3D = ro(pi * integ([&](int k) { return us2[k] * beta[k] * beta[k] * r[k] * std::pow(hg[k], 3) / 12.0; }));
        const double upa_3D = ro(pi * integ([&](int k) { return upa[k] * r[k] * hg[k]; }));
        ts->ekin_2d[0] = us2_2D; ts->ekin_2d[1] = up2_2D; ts->ekin_2d[2] = upa_2D;
        ts->ekin_3d[0] = us2_3D; ts->ekin_3d[1] = up2_3D; ts->ekin_3d[2] = uz2_3D; ts->ekin_3d[3] = upa_3D;
        // Reynolds numbers (:572-584); the sphere container (pre_calculations.f90:50-59)
        const double surf = pi * (h->r_cmb * h->r_cmb - h->r_icb * h->r_icb);
        const double vol_otc = 4.0 / 3.0 * pi * std::pow(h->r_cmb * h->r_cmb - h->r_icb * h->r_icb, 1.5);
        ts->reynolds_2d[0] = std::sqrt(2.0 * (us2_2D + up2_2D) / surf);
        ts->reynolds_2d[1] = std::sqrt(2.0 * upa_2D / surf);
        ts->reynolds_2d[2] = std::sqrt(2.0 * (us2_2D + up2_2D - upa_2D) / surf);
        ts->reynolds_3d[0] = std::sqrt(2.0 * (us2_3D + up2_3D + uz2_3D) / vol_otc);
        ts->reynolds_3d[1] = std::sqrt(2.0 * upa_3D / vol_otc);
        ts->reynolds_3d[2] = std::sqrt(2.0 * (us2_3D + up2_3D + uz2_3D - upa_3D) / vol_otc);
        // power budget (:586-670)
        double visc_2D = ro(2.0 * pi * integ([&](int k) { return enst[k] * r[k]; }));
        if (h->c.kbotv == 1) visc_2D += 4.0 * pi * up2[N - 1];
        if (h->c.ktopv == 1) visc_2D -= 4.0 * pi * up2[0];
        const double pow_2D = h->l_heat ? ro(2.0 * pi * integ([&](int k) { return BuoFac * buo[k] * rc[k] * r[k]; })) : 0.0;
        const double chem_2D = h->l_chem ? ro(2.0 * pi * integ([&](int k) { return ChemFac * chem[k] * rc[k] * r[k]; })) : 0.0;
        double visc_3D = ro(2.0 * pi * integ([&](int k) { return enst[k] * r[k] * hg[k]; }));
        if (h->c.kbotv == 1) visc_3D -= 4.0 * h->or1[N - 1] * up2[N - 1];
        if (h->c.ktopv == 1) visc_3D -= 4.0 * h->or1[0] * up2[0];
        const double pow_3D = h->l_heat ? ro(2.0 * pi * integ([&](int k) { return BuoFac * buo[k] * rc[k] * r[k] * hg[k]; })) : 0.0;
        const double chem_3D = h->l_chem ? ro(2.0 * pi * integ([&](int k) { return ChemFac * chem[k] * rc[k] * r[k] * hg[k]; })) : 0.0;
        const double ptot_2D = ro(2.0 * pi * integ([&](int k) { return CorFac * (us2[k] + up2[k]) * ekp[k] * r[k]; }));
        const double pzon_2D = ro(2.0 * pi * integ([&](int k) { return CorFac * upa[k] * ekp[k] * r[k]; }));
        const double ptot_3D = ro(2.0 * pi * integ([&](int k) { return CorFac * (us2[k] + up2[k]) * ekp[k] * hg[k] * r[k]; }));
        const double pzon_3D = ro(2.0 * pi * integ([&](int k) { return CorFac * upa[k] * ekp[k] * hg[k] * r[k]; }));
        ts->power_2d[0] = pow_2D; ts->power_2d[1] = chem_2D; ts->power_2d[2] = visc_2D;
        ts->power_2d[3] = pzon_2D; ts->power_2d[4] = ptot_2D;
        ts->power_3d[0] = pow_3D; ts->power_3d[1] = chem_3D; ts->power_3d[2] = visc_3D;
        ts->power_3d[3] = pzon_3D; ts->power_3d[4] = ptot_3D;
        // length scales (get_lengthscales, :897-962; dl_diss, :694-701)
        std::vector<double> us2m(nm), up2m(nm), enm(nm), tm(nm), xm(nm);
        for (int i = 0; i < nm; i++) {
            double a[5] = {0, 0, 0, 0, 0};
            for (int j = 0; j < 5; j++)
                for (int k = 0; k < N; k++)
                    a[j] += rint_weights(h)[k] * (pi * spec[(size_t)j * F + (size_t)i * N + k] * r[k] * hg[k]);
            us2m[i] = a[0]; up2m[i] = a[1]; enm[i] = a[2]; tm[i] = a[3]; xm[i] = a[4];
        }
        auto peak = [&](const std::vector<double> &v) {
            int best = -1;
            for (int i = 0; i < nm; i++)
                if (h->idx2m[i] != 0 && (best < 0 || v[i] > v[best])) best = i;
            return best < 0 ? 0 : h->idx2m[best];
        };
        std::vector<double> ek(nm);
        for (int i = 0; i < nm; i++) ek[i] = us2m[i] + up2m[i];
        const int lus = peak(us2m), lek = peak(ek), lvo = peak(enm);
        ts->lengths[0] = lus > 0 ? pi / lus : 0.0;
        ts->lengths[1] = lek > 0 ? pi / lek : 0.0;
        ts->lengths[2] = lvo > 0 ? pi / lvo : 0.0;
        double E = 0.0, Em = 0.0;
        for (int i = 0; i < nm; i++) { E += ek[i]; Em += (double)h->idx2m[i] * ek[i]; }
        ts->lengths[3] = std::fabs(E) > tiny ? pi * E / Em : 0.0;
        // a non-rotating run takes the 2D dissipation length (:697-709)
        if (h->c.l_non_rot)
            ts->lengths[4] = std::fabs(visc_2D) > tiny ? std::sqrt(2.0 * (us2_2D + up2_2D) * ViscFac / visc_2D) : 0.0;
        else
            ts->lengths[4] = std::fabs(visc_3D) > tiny ? std::sqrt(2.0 * (us2_3D + up2_3D) * ViscFac / visc_3D) : 0.0;
        // heat and composition (:720-828), m = 0 at index 0
        const int i0 = 0, mid = N / 2 - 1; // int(n_r_max/2), 1-based
        if (h->l_heat) {
            ts->heat[4] = ro(T[(size_t)i0 * N].x + tcond[0]);
            ts->heat[5] = ro(T[(size_t)i0 * N + N - 1].x + tcond[N - 1]);
            ts->heat[0] = ro(std::fabs(dtcond[0]) > tiny ? 1.0 + dT[(size_t)i0 * N].x / dtcond[0] : 1.0);
            ts->heat[1] = ro(std::fabs(dtcond[N - 1]) > tiny ? 1.0 + dT[(size_t)i0 * N + N - 1].x / dtcond[N - 1] : 1.0);
            for (int k = 0; k < N; k++) nuv[k] *= r[k] * hg[k];
            const double nv = rint_w(h, nuv);
            const double nc = integ([&](int k) { return dtcond[k] * dtcond[k] * hg[k] * r[k]; });
            ts->heat[2] = ro(1.0 + nv / nc);
            std::vector<double> th(N), f1(N), f2(N);
            for (int k = 0; k < N; k++) {
                const double sr = r[k] / h->r_cmb;
                f1[k] = -TdiffFac * dtcond[k] * sr * sr;
                th[k] = std::asin(sr);
                f2[k] = flux[k] * sr * sr;
            }
            if (h->c.l_non_rot) {
                // :763-767
                const double fac = integ([&](int k) { return -TdiffFac * dtcond[k] * r[k]; });
                ts->heat[3] = ro(integ([&](int k) { return flux[k] * r[k] / fac; }));
            } else {
                ts->heat[3] = ro(simps_x(f2, th) / simps_x(f1, th));
            }
            ts->heat[6] = ro(dtcond[mid] + dT[(size_t)i0 * N + mid].x);
        }
        if (h->l_chem) {
            ts->chem[3] = ro(X[(size_t)i0 * N].x + xicond[0]);
            ts->chem[4] = ro(X[(size_t)i0 * N + N - 1].x + xicond[N - 1]);
            ts->chem[0] = ro(1.0 + dX[(size_t)i0 * N].x / dxicond[0]);
            ts->chem[1] = ro(1.0 + dX[(size_t)i0 * N + N - 1].x / dxicond[N - 1]);
            for (int k = 0; k < N; k++) shv[k] *= r[k] * hg[k];
            const double sv = rint_w(h, shv);
            const double sc = integ([&](int k) { return dxicond[k] * dxicond[k] * hg[k] * r[k]; });
            ts->chem[2] = ro(1.0 + sv / sc);
            ts->chem[5] = ro(dxicond[mid] + dX[(size_t)i0 * N + mid].x);
        }
        if (h->c.l_non_rot) {
            // a non-rotating run writes no 3D series (outputs.f90:102-121)
            for (double &v : ts->ekin_3d) v = 0.0;
            for (double &v : ts->power_3d) v = 0.0;
            for (double &v : ts->reynolds_3d) v = 0.0;
        }
        // the m-spectra of the same call (calculate_spectra): us2, up2,
        // enstrophy, T^2 and xi^2 per m, kept for pz_spectra
        h->spec.assign(5 * (size_t)nm, 0.0);
        for (int i = 0; i < nm; i++) {
            h->spec[i] = us2m[i];
            h->spec[nm + i] = up2m[i];
            h->spec[2 * nm + i] = enm[i];
            h->spec[3 * nm + i] = tm[i];
            // spectra.f90:188 integrates the temperature profile for xi2_m
            h->spec[4 * nm + i] = h->l_heat ? tm[i] : xm[i];
        }
    } catch (std::exception &e) {
        return fail(h, -4, e.what());
    }
    return 0;
}

extern "C" int pz_spectra(pz_handle *h, double *us2_m, double *up2_m, double *enst_m, double *temp2_m, double *xi2_m)
{
    if (!h) return -1;
    pz_series ts;
    const int rc = pz_time_series(h, &ts);
    if (rc) return rc;
    double *outs[5] = {us2_m, up2_m, enst_m, temp2_m, xi2_m};
    for (int j = 0; j < 5; j++)
        if (outs[j]) std::memcpy(outs[j], h->spec.data() + (size_t)j * h->nm, sizeof(double) * h->nm);
    return 0;
}

// ---- unit operators on caller host buffers --------------------------------
static int vec_op(pz_handle *h, const double *f, double *df, double *ddf, int n_m, int op, int l_dct_in)
{
    try {
        const int N = h->N;
        const size_t S = (size_t)n_m * N;
        std::vector<double2> v(S);
        for (int k = 0; k < N; k++)
            for (int i = 0; i < n_m; i++) {
                size_t s = 2 * ((size_t)k * n_m + i);
                v[(size_t)i * N + k] = make_double2(f[s], f[s + 1]);
            }
        cd *a = nullptr, *b = nullptr, *c = nullptr;
        HIPCHK(hipMalloc(&a, S * sizeof(cd)));
        HIPCHK(hipMalloc(&b, S * sizeof(cd)));
        HIPCHK(hipMalloc(&c, S * sizeof(cd)));
        HIPCHK(hipMemcpy(a, v.data(), S * sizeof(cd), hipMemcpyHostToDevice));
        if (op == 0) hipLaunchKernelGGL(k_costf_vecs, dim3(n_m), dim3(256), (2 * N + 2 * (size_t)h->d.L) * sizeof(cd), h->stream, h->d, a);
        else hipLaunchKernelGGL(k_dr_vecs, dim3(n_m), dim3(256), (4 * N + 4 * (size_t)h->d.L + 16) * sizeof(cd), h->stream, h->d, a, b,
                                op == 2 ? c : nullptr, l_dct_in);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(h->stream));
        auto back = [&](cd *src, double *dst) {
            HIPCHK(hipMemcpy(v.data(), src, S * sizeof(cd), hipMemcpyDeviceToHost));
            for (int k = 0; k < N; k++)
                for (int i = 0; i < n_m; i++) {
                    size_t s = 2 * ((size_t)k * n_m + i);
                    dst[s] = v[(size_t)i * N + k].x;
                    dst[s + 1] = v[(size_t)i * N + k].y;
                }
        };
        if (op == 0) back(a, df);
        else {
            back(b, df);
            if (op == 2) back(c, ddf);
        }
        (void)hipFree(a);
        (void)hipFree(b);
        (void)hipFree(c);
    } catch (std::exception &e) {
        return fail(h, -4, e.what());
    }
    return 0;
}

extern "C" int pz_costf(pz_handle *h, double *f, int n_m)
{
    if (!h || !f || n_m <= 0) return -1;
    return vec_op(h, f, f, nullptr, n_m, 0, 0);
}
extern "C" int pz_get_dr(pz_handle *h, const double *f, double *df, int n_m, int l_dct_in)
{
    if (!h || !f || !df || n_m <= 0) return -1;
    return vec_op(h, f, df, nullptr, n_m, 1, l_dct_in);
}
extern "C" int pz_get_ddr(pz_handle *h, const double *f, double *df, double *ddf, int n_m, int l_dct_in)
{
    if (!h || !f || !df || !ddf || n_m <= 0) return -1;
    return vec_op(h, f, df, ddf, n_m, 2, l_dct_in);
}

// get_dr / get_ddr on device buffers in the reference layout, staged through
// the static scratch array: no allocation, asynchronous on the handle's
// stream.  The transform kernel is pz_get_ddr's.
extern "C" int pz_get_ddr_noalloc(pz_handle *h, const double *d_f, double *d_df, double *d_ddf, int n_m,
                                  int l_dct_in)
{
    if (!h || !d_f || !d_df || n_m <= 0) return -1;
    const int N = h->N;
    const size_t S = (size_t)n_m * N;
    if (S > PZ_SCRATCH_ELEMS)
        return fail(h, -2, "pz_get_ddr_noalloc: n_m * n_r_max exceeds the staging limit of 262144 complex values");
    if (!h->scratch) return fail(h, -4, "pz_get_ddr_noalloc: no staging address on this device");
    double2 *a = h->scratch, *b = a + PZ_SCRATCH_ELEMS, *c = b + PZ_SCRATCH_ELEMS;
    const int blocks = (int)std::min<size_t>((S + 255) / 256, 1024);
    hipLaunchKernelGGL(k_from_ref_layout, dim3(blocks), dim3(256), 0, h->stream, (const double2 *)d_f, a, n_m, N);
    hipLaunchKernelGGL(k_dr_vecs, dim3(n_m), dim3(256), (4 * N + 4 * (size_t)h->d.L + 16) * sizeof(cd), h->stream, h->d,
                       a, b, d_ddf ? c : nullptr, l_dct_in);
    hipLaunchKernelGGL(k_to_ref_layout, dim3(blocks), dim3(256), 0, h->stream, b, (double2 *)d_df, n_m, N);
    if (d_ddf) hipLaunchKernelGGL(k_to_ref_layout, dim3(blocks), dim3(256), 0, h->stream, c, (double2 *)d_ddf, n_m, N);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(h, -4, std::string("pz_get_ddr_noalloc: ") + hipGetErrorString(e));
    return 0;
}

// Classical top and bottom Nusselt numbers of the heat time series
// (outputs.f90:729-755): 1 + (dT/dr of the m = 0 perturbation) / (dT_cond/dr)
// at each wall, from the radial derivative of the run's own method.
extern "C" int pz_nusselt(pz_handle *h, double out[2])
{
    if (!h || !out) return -1;
    const double tiny = 10.0 * 2.220446049250313e-16;
    if (!h->l_heat) {
        out[0] = out[1] = 1.0;
        return 0;
    }
    const int N = h->N, nm = h->nm;
    std::vector<double> fld(2 * (size_t)nm * N), col(2 * (size_t)N), dcol(2 * (size_t)N);
    int rc;
    if (h->d.fd) {
        // the finite-difference step refreshes dT/dr (ghost-zone stencil) every stage
        rc = pz_get_field(h, PZ_DTEMP, fld.data());
        if (rc) return rc;
        for (int k = 0; k < N; k++) dcol[2 * k] = fld[2 * ((size_t)k * nm)];
    } else {
        rc = pz_get_field(h, PZ_TEMP, fld.data());
        if (rc) return rc;
        for (int k = 0; k < N; k++) { col[2 * k] = fld[2 * ((size_t)k * nm)]; col[2 * k + 1] = fld[2 * ((size_t)k * nm) + 1]; }
        rc = pz_get_dr(h, col.data(), dcol.data(), 1, 1);
        if (rc) return rc;
    }
    auto round_off = [&](double v) { return std::fabs(v) < tiny ? 0.0 : v; };
    out[0] = round_off(std::fabs(h->dtcond_top) > tiny ? 1.0 + dcol[0] / h->dtcond_top : 1.0);
    out[1] = round_off(std::fabs(h->dtcond_bot) > tiny ? 1.0 + dcol[2 * (N - 1)] / h->dtcond_bot : 1.0);
    return 0;
}

// ---------------------------------------------------------------------------
// Adjoint of the step (no reference counterpart).  The forward sweep keeps a
// snapshot of the step state every ck steps; the backward sweep replays each
// segment from its snapshot, keeping the state before every step of it,
// and runs the transposed step kernels backwards through the segment.
// ---------------------------------------------------------------------------
// the implicit systems a step reads (pointer, bytes): the collocation
// inverses, or the factorised finite-difference bands
static std::vector<std::pair<double *, size_t>> adj_live(pz_handle *h)
{
    const size_t nm = h->nm, N = h->N;
    if (h->d.fd)
        return {{h->d.fT, nm * 3 * N * sizeof(double)}, {h->d.fX, nm * 3 * N * sizeof(double)},
                {h->d.fU, 3 * N * sizeof(double)}, {(double *)h->d.fP, nm * 5 * N * sizeof(cd)}};
    const size_t ti = nm * N * h->d.ldi, pc = nm * N * 2 * N, ui = N * h->d.ldi;
    return {{h->d.T_inv, ti * sizeof(double)}, {h->d.X_inv, ti * sizeof(double)}, {h->d.U_inv, ui * sizeof(double)},
            {(double *)h->d.PC_inv, pc * sizeof(cd)}};
}

static void adj_init(pz_handle *h)
{
    if (h->adj) return;
    AdjCtx *a = new AdjCtx();
    h->adj = a;
    const int N = h->N, nm = h->nm;
    const size_t F = (size_t)nm * N, NN = (size_t)N * N;
    PzAdjDev &ad = a->ad;
    double *C = h->dalloc<double>(NN), *G = h->dalloc<double>(NN), *D1 = h->dalloc<double>(NN);
    double *D2 = h->dalloc<double>(NN), *E1 = h->dalloc<double>(NN), *U = h->dalloc<double>(NN);
    double *lnu = h->dalloc<double>(N);
    ad.C = C; ad.G = G; ad.D1 = D1; ad.D2 = D2; ad.E1 = E1; ad.Uimp = U; ad.lnu = lnu;
    {
        // kinetic energy weights (outputs.f90:552-570): rInt_R is linear, so
        // its quadrature weights are its values on the unit vectors
        const double pi = 4.0 * std::atan(1.0);
        std::vector<double> kus(N), kup(N), e(N, 0.0);
        for (int k = 0; k < N; k++) {
            e[k] = 1.0;
            const double q = pi * rint_r(h, e), rh = h->r[k] * h->height[k];
            e[k] = 0.0;
            kup[k] = q * rh;
            kus[k] = q * (rh + h->beta[k] * h->beta[k] * h->r[k] * std::pow(h->height[k], 3) / 12.0);
        }
        ad.kus = h->upload(kus);
        ad.kup = h->upload(kup);
    }
    ad.eT0 = h->dalloc<cd>(F); ad.aT0 = h->dalloc<cd>(F); ad.eX0 = h->dalloc<cd>(F);
    ad.aX0 = h->dalloc<cd>(F); ad.eP0 = h->dalloc<cd>(F); ad.aP0 = h->dalloc<cd>(F);
    ad.gtop = h->dalloc<cd>(nm); ad.gbot = h->dalloc<cd>(nm);
    ad.gtopxi = h->dalloc<cd>(nm); ad.gbotxi = h->dalloc<cd>(nm);
    ad.flag = h->dalloc<int>(1);
    {
        // copies of the implicit systems (the inverses of the collocation
        // method, the factorised bands of finite differences): S = those of
        // the scheme's own weights, B = those of the CNAB2 start-up steps of
        // a bridged scheme
        a->mat_bytes.clear();
        for (auto &lv : adj_live(h)) a->mat_bytes.push_back(lv.second);
        auto alloc_set = [&]() {
            std::vector<double *> v;
            for (size_t b : a->mat_bytes) v.push_back((double *)h->dalloc<char>(b));
            return v;
        };
        a->mats_s = alloc_set();
        a->mats_b = h->d.bridge ? alloc_set() : a->mats_s;
        if (h->d.fd) {
            ad.fT_s = a->mats_s[0]; ad.fX_s = a->mats_s[1]; ad.fU_s = a->mats_s[2]; ad.fP_s = (const cd *)a->mats_s[3];
            ad.fT_b = a->mats_b[0]; ad.fX_b = a->mats_b[1]; ad.fU_b = a->mats_b[2]; ad.fP_b = (const cd *)a->mats_b[3];
        } else {
            ad.T_inv_s = a->mats_s[0]; ad.X_inv_s = a->mats_s[1]; ad.U_inv_s = a->mats_s[2];
            ad.PC_inv_s = (const cd *)a->mats_s[3];
            ad.T_inv_b = a->mats_b[0]; ad.X_inv_b = a->mats_b[1]; ad.U_inv_b = a->mats_b[2];
            ad.PC_inv_b = (const cd *)a->mats_b[3];
        }
    }
    HIPCHK(hipStreamCreate(&a->s2));
    HIPCHK(hipEventCreateWithFlags(&a->ev[0], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&a->ev[1], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&a->ev[2], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&a->ev[3], hipEventDisableTiming));
    if (h->radial_inplace) ad.gfs = h->dalloc<cd>(2 * (size_t)h->nphi * h->N);
    a->dad = h->upload(std::vector<PzAdjDev>{ad});
    a->lam = h->dalloc<char>(h->arena_bytes);
    PzDev lv = h->d;
    arena_layout(lv, a->lam, F, h->sch.nold, h->sch.nimp, h->sch.nexp);
    lv.ss = h->d.ss;
    a->dlam = h->upload(std::vector<PzDev>{lv});
    a->acc = h->dalloc<double>(1 + (size_t)h->d.nm); // total, then one partial sum per m
    a->W = h->dalloc<cd>(F);
    a->grad = h->dalloc<cd>(F);
    if (h->radial_inplace) a->rrow = h->dalloc<cd>(4 * F);
    if (h->d.fd) {
        // dT/dr at the outer wall as the step takes it (the ghost-point
        // stencil of row 0, fd_implicit): a_1 T_0 + a_2 T_1 + a_0 T_-1 with
        // T_-1 = 2 T_0 - T_1 (fixed value) or T_1 (fixed flux)
        std::vector<double> row(N, 0.0), fdc(3);
        HIPCHK(hipMemcpy(fdc.data(), h->d.fdc, 3 * sizeof(double), hipMemcpyDeviceToHost));
        if (h->c.ktopt == 1) { row[0] = fdc[1] + 2.0 * fdc[0]; row[1] = fdc[2] - fdc[0]; }
        else { row[0] = fdc[1]; row[1] = fdc[2] + fdc[0]; }
        HIPCHK(hipMemcpy(lnu, row.data(), N * sizeof(double), hipMemcpyHostToDevice));
        return;
    }
    const size_t lds = (6 * (size_t)N + 6 * (size_t)h->d.L + 16) * sizeof(cd) + 3 * (size_t)N * sizeof(double);
    hipLaunchKernelGGL(k_adj_probe, dim3(N), dim3(256), lds, h->stream, (const PzDev *)h->ddev, C, G, D1, D2, E1, U,
                       lnu);
    HIPCHK(hipGetLastError());
}

// snapshot blocks and views for a window of K steps, checkpoints every ck
static void adj_window(pz_handle *h, int K, int ck)
{
    AdjCtx *a = h->adj;
    const size_t A = h->arena_bytes, F = (size_t)h->nm * h->N;
    const int nseg = (K + ck - 1) / ck;
    if (nseg > a->cap_ckpt || ck > a->cap_ck) {
        a->release();
        HIPCHK(hipMalloc(&a->ckpt, A * nseg));
        HIPCHK(hipMalloc(&a->seg, A * 2 * (ck + 1)));
        HIPCHK(hipMalloc(&a->fin, A));
        HIPCHK(hipMalloc(&a->dviews, sizeof(PzDev) * 2 * ck));
        a->cap_ckpt = nseg;
        a->cap_ck = ck;
        a->views_ck = 0;
    }
    if (a->views_ck != ck) {
        // two segment buffers of ck + 1 snapshots; view j of buffer b: the
        // state before step j of the segment, with the scheme state the
        // step ran with (the one stored before step j + 1).  The layout
        // follows ck, so the views are rebuilt whenever ck changes.
        std::vector<PzDev> v(2 * ck, h->d);
        for (int b = 0; b < 2; b++)
            for (int j = 0; j < ck; j++) {
                PzDev next = h->d;
                char *base = a->seg + A * (b * (ck + 1) + j);
                arena_layout(v[b * ck + j], base, F, h->sch.nold, h->sch.nimp, h->sch.nexp);
                arena_layout(next, base + A, F, h->sch.nold, h->sch.nimp, h->sch.nexp);
                v[b * ck + j].ss = next.ss;
                if (a->rrow) v[b * ck + j].rrow = a->rrow; // apart from the replay's own
            }
        HIPCHK(hipMemcpy(a->dviews, v.data(), sizeof(PzDev) * 2 * ck, hipMemcpyHostToDevice));
        a->views_ck = ck;
    }
    a->K = K;
    a->ck = ck;
    a->nseg = nseg;
}

// copies the live inverses to a set (to_live = false) or back
static void adj_mats(pz_handle *h, const std::vector<double *> &set, bool to_live)
{
    const auto live = adj_live(h);
    for (size_t i = 0; i < live.size(); i++)
        if (live[i].first && live[i].second)
            HIPCHK(hipMemcpyAsync(to_live ? live[i].first : set[i], to_live ? set[i] : live[i].first,
                                  h->adj->mat_bytes[i], hipMemcpyDeviceToDevice, h->stream));
}

static int adj_check(pz_handle *h)
{
    if (!h->started) return fail(h, -5, "pz_start has not been called");
    if (!(h->d.coll || h->d.fd) || h->sch.family != 0)
        return fail(h, -2, "the adjoint covers the collocation and finite-difference methods with a multistep "
                           "time scheme");
    if (h->d.coll && h->c.l_rerror_fix)
        return fail(h, -2, "the adjoint needs l_rerror_fix = .false. (thresholds are not linear)");
    if (h->radial_big && !h->radial_inplace)
        return fail(h, -2, "the adjoint of the long azimuthal grids needs the run without composition");
    if (!h->l_heat) return fail(h, -2, "the adjoint objectives need the temperature equation");
    if (h->nphi < 2 * h->nm) return fail(h, -2, "the adjoint needs n_phi >= 2 n_m");
    if (h->d.dmat) return fail(h, -2, "the adjoint covers matrix_solve='direct'");
    return 0;
}

extern "C" int pz_adjoint_forward(pz_handle *h, int objective, const double *weights, int nsteps, int ckpt_every,
                                  double *J)
{
    if (!h || !J || nsteps < 1 || ckpt_every < 1) return -1;
    if (objective != PZ_OBJ_NUTOP && objective != PZ_OBJ_EKIN && objective != PZ_OBJ_TEMP)
        return fail(h, -1, "unknown objective");
    if (objective == PZ_OBJ_TEMP && !weights) return fail(h, -1, "the temperature objective needs weights");
    int rc = adj_check(h);
    if (rc) return rc;
    try {
        PzScheme s0;
        HIPCHK(hipMemcpyAsync(&s0, h->d.ss, sizeof(s0), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        if (s0.n_steps != 0) return fail(h, -5, "the adjoint window starts right after pz_start");
        const size_t lds_radj = h->radial_inplace
                                    ? (2 * (size_t)h->nphi + 64 + (h->nphi + 63) / 64 + PZ_FFT_OM_MAX) * sizeof(cd)
                                    : (11 * (size_t)h->nm + 9 * (size_t)h->nphi) * sizeof(cd);
        const size_t lds_madj = h->d.fd ? (3 * ((size_t)h->N + 4) + 10 * (size_t)h->N) * sizeof(cd)
                                        : ADJ_NVEC * (size_t)h->N * sizeof(cd);
        int dev = 0, lim = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        if (lds_radj > (size_t)lim || lds_madj > (size_t)lim)
            return fail(h, -2, "the adjoint kernels need more LDS than a workgroup has at this size");
        allow_full_lds(lim);
        adj_init(h);
        adj_window(h, nsteps, ckpt_every);
        AdjCtx *a = h->adj;
        a->objective = objective;
        a->ready = false;
        const size_t A = h->arena_bytes;
        HIPCHK(hipMemsetAsync(a->acc, 0, sizeof(double), h->stream));
        if (objective == PZ_OBJ_TEMP) {
            std::vector<double2> w;
            to_device_layout(h, weights, w);
            HIPCHK(hipMemcpyAsync(a->W, w.data(), w.size() * sizeof(double2), hipMemcpyHostToDevice, h->stream));
        }
        for (int sg = 0; sg < a->nseg; sg++) {
            HIPCHK(hipMemcpyAsync(a->ckpt + A * sg, h->arena, A, hipMemcpyDeviceToDevice, h->stream));
            const int len = std::min(a->ck, nsteps - sg * a->ck);
            for (int i = 0; i < len; i++) {
                run_steps(h, 1);
                if (h->d.bridge && sg * a->ck + i == 0) adj_mats(h, a->mats_b, false);
                if (objective == PZ_OBJ_NUTOP)
                    hipLaunchKernelGGL(k_obj_acc, dim3(1), dim3(256), 0, h->stream, (const PzDev *)h->ddev,
                                       (const PzAdjDev *)a->dad, (const cd *)nullptr, a->acc);
                else if (objective == PZ_OBJ_EKIN) {
                    hipLaunchKernelGGL(k_obj_ke, dim3(h->d.nm), dim3(256), 0, h->stream, (const PzDev *)h->ddev,
                                       (const PzAdjDev *)a->dad, a->acc);
                    hipLaunchKernelGGL(k_obj_fin, dim3(1), dim3(256), 0, h->stream, h->d.nm, a->acc);
                }
            }
        }
        if (objective == PZ_OBJ_TEMP)
            hipLaunchKernelGGL(k_obj_acc, dim3(1), dim3(256), 0, h->stream, (const PzDev *)h->ddev,
                               (const PzAdjDev *)a->dad, (const cd *)a->W, a->acc);
        HIPCHK(hipMemcpyAsync(a->fin, h->arena, A, hipMemcpyDeviceToDevice, h->stream));
        adj_mats(h, a->mats_s, false);
        HIPCHK(hipGetLastError());
        double acc = 0.0;
        HIPCHK(hipMemcpyAsync(&acc, a->acc, sizeof(double), hipMemcpyDeviceToHost, h->stream));
        rc = check_state(h, nullptr);
        if (rc) return rc;
        const double tiny = 10.0 * 2.220446049250313e-16;
        if (objective == PZ_OBJ_NUTOP)
            *J = std::fabs(h->dtcond_top) > tiny ? 1.0 + acc / ((double)nsteps * h->dtcond_top) : 1.0;
        else if (objective == PZ_OBJ_EKIN)
            *J = acc / (double)nsteps;
        else
            *J = acc;
        a->ready = true;
    } catch (std::exception &e) {
        return fail(h, -4, e.what());
    }
    return 0;
}

extern "C" int pz_adjoint_backward(pz_handle *h, double dJ, double *grad_temp, double *grad_bc)
{
    if (!h) return -1;
    if (!h->adj || !h->adj->ready) return fail(h, -5, "pz_adjoint_backward needs a pz_adjoint_forward first");
    try {
        AdjCtx *a = h->adj;
        const size_t A = h->arena_bytes, F = (size_t)h->nm * h->N;
        const int K = a->K, ck = a->ck, nm = h->nm, N = h->N;
        const size_t lds_radj = (11 * (size_t)nm + 9 * (size_t)h->nphi) * sizeof(cd);
        const size_t lds_madj = ADJ_NVEC * (size_t)N * sizeof(cd);
        const size_t lds_fadj = (3 * ((size_t)N + 4) + 10 * (size_t)N) * sizeof(cd);
        const size_t lds_radj_in = (2 * (size_t)h->nphi + 64 + (h->nphi + 63) / 64 + PZ_FFT_OM_MAX) * sizeof(cd);
        const double tiny = 10.0 * 2.220446049250313e-16;
        double wnu = 0.0;
        if (a->objective == PZ_OBJ_NUTOP && std::fabs(h->dtcond_top) > tiny) wnu = dJ / ((double)K * h->dtcond_top);
        if (a->objective == PZ_OBJ_EKIN) wnu = dJ / (double)K;
        // u_s and u_phi of a snapshot sit at fixed offsets of its block
        const size_t off_us = (size_t)((char *)h->d.us - h->arena), off_up = (size_t)((char *)h->d.up - h->arena);
        HIPCHK(hipMemsetAsync(a->lam, 0, A, h->stream));
        for (cd *q : {a->ad.gtop, a->ad.gbot, a->ad.gtopxi, a->ad.gbotxi})
            HIPCHK(hipMemsetAsync(q, 0, nm * sizeof(cd), h->stream));
        HIPCHK(hipMemsetAsync(a->ad.flag, 0, sizeof(int), h->stream));
        if (a->objective == PZ_OBJ_TEMP && dJ != 1.0) {
            // dJ W as the seed
            std::vector<double2> w(F);
            HIPCHK(hipMemcpyAsync(w.data(), a->W, F * sizeof(double2), hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            for (auto &v : w) { v.x *= dJ; v.y *= dJ; }
            HIPCHK(hipMemcpyAsync(a->grad, w.data(), F * sizeof(double2), hipMemcpyHostToDevice, h->stream));
        } else if (a->objective == PZ_OBJ_TEMP) {
            HIPCHK(hipMemcpyAsync(a->grad, a->W, F * sizeof(double2), hipMemcpyDeviceToDevice, h->stream));
        }
        // replay of segment sg into buffer sg % 2 on the handle's stream
        auto replay = [&](int sg) {
            const int len = std::min(ck, K - sg * ck), c = sg * ck, b = sg % 2;
            char *seg = a->seg + A * (size_t)(b * (ck + 1));
            HIPCHK(hipMemcpyAsync(h->arena, a->ckpt + A * sg, A, hipMemcpyDeviceToDevice, h->stream));
            // a replay from inside the start-up needs its matrices (the
            // first step and the end of the start-up rebuild them)
            if (h->d.bridge && c >= 1 && c <= h->sch.nold - 2) adj_mats(h, a->mats_b, true);
            if (h->d.bridge && c >= h->sch.nold) adj_mats(h, a->mats_s, true);
            for (int j = 0; j < len; j++) {
                HIPCHK(hipMemcpyAsync(seg + A * j, h->arena, A, hipMemcpyDeviceToDevice, h->stream));
                run_steps(h, 1);
            }
            HIPCHK(hipMemcpyAsync(seg + A * len, h->arena, A, hipMemcpyDeviceToDevice, h->stream));
            HIPCHK(hipEventRecord(a->ev[b], h->stream));
        };
        // transposed steps of segment sg on s2, last step first
        auto adjoint = [&](int sg) {
            const int len = std::min(ck, K - sg * ck), b = sg % 2;
            HIPCHK(hipStreamWaitEvent(a->s2, a->ev[b], 0));
            for (int j = len - 1; j >= 0; j--) {
                const int n = sg * ck + j + 1; // the state after step n seeds the objective
                const PzDev *fw = a->dviews + b * ck + j;
                if (a->objective == PZ_OBJ_NUTOP) {
                    hipLaunchKernelGGL(k_adj_seed_nu, dim3(1), dim3(256), 0, a->s2, (const PzDev *)a->dlam,
                                       (const PzAdjDev *)a->dad, wnu);
                } else if (a->objective == PZ_OBJ_EKIN) {
                    // the state after step n: snapshot j + 1 of the segment buffer
                    const char *after = a->seg + A * (size_t)(b * (ck + 1) + j + 1);
                    hipLaunchKernelGGL(k_adj_seed_ke, dim3((unsigned)((F + 255) / 256)), dim3(256), 0, a->s2,
                                       (const cd *)(after + off_us), (const cd *)(after + off_up),
                                       (const PzDev *)a->dlam, (const PzAdjDev *)a->dad, wnu);
                } else if (n == K)
                    hipLaunchKernelGGL(k_adj_seed_w, dim3((unsigned)((F + 255) / 256)), dim3(256), 0, a->s2,
                                       (const PzDev *)a->dlam, (const cd *)a->grad);
                if (h->d.fd)
                    hipLaunchKernelGGL(k_fd_mloop_adj, dim3(nm), dim3(256), lds_fadj, a->s2, fw,
                                       (const PzDev *)a->dlam, (const PzAdjDev *)a->dad);
                else
                    hipLaunchKernelGGL(k_mcoll_adj, dim3(nm), dim3(512), lds_madj, a->s2, fw, (const PzDev *)a->dlam,
                                       (const PzAdjDev *)a->dad);
                if (h->radial_inplace) {
                    hipLaunchKernelGGL(k_gather_r,
                                       dim3((nm + PZ_GATHER_TM - 1) / PZ_GATHER_TM, (N + PZ_GATHER_TR - 1) / PZ_GATHER_TR),
                                       dim3(256), 0, a->s2, fw);
                    hipLaunchKernelGGL(k_radial_adj_inplace, dim3(N), dim3(PZ_RADIAL_ADJ_THREADS), lds_radj_in,
                                       a->s2, fw, (const PzDev *)a->dlam, (const PzAdjDev *)a->dad);
                } else {
                    hipLaunchKernelGGL(k_radial_adj, dim3(N), dim3(256), lds_radj, a->s2, fw, (const PzDev *)a->dlam,
                                       (const PzAdjDev *)a->dad);
                }
            }
            HIPCHK(hipEventRecord(a->ev[2 + b], a->s2));
        };
        // the adjoint variables and seeds are set on the handle's stream
        HIPCHK(hipEventRecord(a->ev[2], h->stream));
        HIPCHK(hipStreamWaitEvent(a->s2, a->ev[2], 0));
        replay(a->nseg - 1);
        for (int sg = a->nseg - 1; sg >= 0; sg--) {
            if (sg >= 1) {
                // buffer (sg - 1) % 2 was last read by the adjoint of sg + 1
                if (sg + 1 < a->nseg) HIPCHK(hipStreamWaitEvent(h->stream, a->ev[2 + (sg + 1) % 2], 0));
                replay(sg - 1);
            }
            adjoint(sg);
        }
        HIPCHK(hipStreamWaitEvent(h->stream, a->ev[2], 0));
        if (h->d.fd)
            hipLaunchKernelGGL(k_fd_start_adj, dim3(nm), dim3(256), (4 * (size_t)N + 4) * sizeof(cd), h->stream,
                               (const PzDev *)h->ddev, (const PzDev *)a->dlam, (const PzAdjDev *)a->dad, a->grad);
        else
            hipLaunchKernelGGL(k_start_coll_adj, dim3(nm), dim3(256), 6 * (size_t)N * sizeof(cd), h->stream,
                               (const PzDev *)h->ddev, (const PzDev *)a->dlam, (const PzAdjDev *)a->dad, a->grad);
        HIPCHK(hipMemcpyAsync(h->arena, a->fin, A, hipMemcpyDeviceToDevice, h->stream));
        adj_mats(h, a->mats_s, true);
        HIPCHK(hipGetLastError());
        int flag = 0;
        std::vector<double2> g(F), bc(2 * (size_t)nm);
        HIPCHK(hipMemcpyAsync(&flag, a->ad.flag, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(g.data(), a->grad, F * sizeof(double2), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(bc.data(), a->ad.gtop, nm * sizeof(double2), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(bc.data() + nm, a->ad.gbot, nm * sizeof(double2), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        if (flag) return fail(h, -7, "dt or the implicit matrices changed inside the adjoint window");
        if (grad_temp) from_device_layout(h, g, grad_temp);
        if (grad_bc) std::memcpy(grad_bc, bc.data(), bc.size() * sizeof(double2));
    } catch (std::exception &e) {
        return fail(h, -4, e.what());
    }
    return 0;
}

extern "C" int pz_fft(pz_handle *h, const double *grid, double *spec, int n_rad)
{
    if (!h || !grid || !spec || n_rad <= 0) return -1;
    try {
        double *g = nullptr;
        cd *s = nullptr;
        HIPCHK(hipMalloc(&g, (size_t)n_rad * h->nphi * sizeof(double)));
        HIPCHK(hipMalloc(&s, (size_t)n_rad * h->nm * sizeof(cd)));
        HIPCHK(hipMemcpy(g, grid, (size_t)n_rad * h->nphi * sizeof(double), hipMemcpyHostToDevice));
        if (h->radial_inplace)
            hipLaunchKernelGGL(k_phi_fft_inplace, dim3(n_rad), dim3(PZ_RADIAL_INPLACE_THREADS), h->lds_radial,
                               h->stream, (const PzDev *)h->ddev, g, s);
        else
            hipLaunchKernelGGL(k_phi_fft, dim3(n_rad), dim3(256), 2 * h->nphi * sizeof(cd), h->stream, h->d, g, s);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipMemcpy(spec, s, (size_t)n_rad * h->nm * sizeof(cd), hipMemcpyDeviceToHost));
        (void)hipFree(g);
        (void)hipFree(s);
    } catch (std::exception &e) {
        return fail(h, -4, e.what());
    }
    return 0;
}

extern "C" int pz_ifft(pz_handle *h, const double *spec, double *grid, int n_rad)
{
    if (!h || !grid || !spec || n_rad <= 0) return -1;
    try {
        double *g = nullptr;
        cd *s = nullptr;
        HIPCHK(hipMalloc(&g, (size_t)n_rad * h->nphi * sizeof(double)));
        HIPCHK(hipMalloc(&s, (size_t)n_rad * h->nm * sizeof(cd)));
        HIPCHK(hipMemcpy(s, spec, (size_t)n_rad * h->nm * sizeof(cd), hipMemcpyHostToDevice));
        if (h->radial_inplace)
            hipLaunchKernelGGL(k_phi_ifft_inplace, dim3(n_rad), dim3(PZ_RADIAL_INPLACE_THREADS), h->lds_radial,
                               h->stream, (const PzDev *)h->ddev, s, g);
        else
            hipLaunchKernelGGL(k_phi_ifft, dim3(n_rad), dim3(256), 2 * h->nphi * sizeof(cd), h->stream, h->d, s, g);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipMemcpy(grid, g, (size_t)n_rad * h->nphi * sizeof(double), hipMemcpyDeviceToHost));
        (void)hipFree(g);
        (void)hipFree(s);
    } catch (std::exception &e) {
        return fail(h, -4, e.what());
    }
    return 0;
}

// which: 0 = m-loop stage kernel (stage 1 of the scheme, exp_calc), 1 = radial loop
extern "C" int pz_time_kernel(pz_handle *h, int which, int reps, double *avg_ms)
{
    if (!h || !avg_ms || reps <= 0) return -1;
    if (!h->started) return fail(h, -5, "pz_start has not been called");
    try {
        // the m-loop kernels advance the fields and the time-scheme slots:
        // keep a copy of the step state and put it back after the timing
        char *keep = nullptr;
        HIPCHK(hipMalloc(&keep, h->arena_bytes));
        HIPCHK(hipMemcpyAsync(keep, h->arena, h->arena_bytes, hipMemcpyDeviceToDevice, h->stream));
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, h->stream));
        for (int i = 0; i < reps; i++) {
            if (which == 0 && h->d.fd)
                hipLaunchKernelGGL(k_fd_mloop, dim3(h->nm), dim3(256), h->lds_fd_loop, h->stream, (const PzDev *)h->ddev, 1,
                                   1);
            else if (which == 0 && h->d.coll)
                hipLaunchKernelGGL(h->mcoll_l2 ? k_mcoll_l2 : k_mcoll, dim3(h->nm), dim3(h->mthreads), h->lds_mphase, h->stream,
                                   (const PzDev *)h->ddev, 1, 1);
            else if (which == 0)
                hipLaunchKernelGGL(k_mphase, dim3(h->nm), dim3(h->mthreads), h->lds_mphase, h->stream,
                                   (const PzDev *)h->ddev, 1, 1);
            else
                launch_radial(h, 0, 0);
        }
        HIPCHK(hipEventRecord(e1, h->stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        *avg_ms = (double)ms / reps;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        HIPCHK(hipMemcpyAsync(h->arena, keep, h->arena_bytes, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipFree(keep));
    } catch (std::exception &e) {
        return fail(h, -4, e.what());
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Ensembles: every kernel of the stage loop is launched once for all
// members, each block finding its run and its own block index in a map.
// ---------------------------------------------------------------------------
// One size class of an ensemble (members of one n_r): its own maps, LDS
// budgets, stream and captured step graph, so that classes run side by
// side and the small members are not given the LDS of the largest.
struct EnsClass {
    std::vector<int> members;
    const PzDev **druns = nullptr;
    int2 *map_r = nullptr, *map_c = nullptr, *map_3 = nullptr, *map_m = nullptr;
    int nb_r = 0, nb_c = 0, nb_3 = 0, nb_m = 0;
    size_t lds_r = 0, lds_m = 0, lds_b = 0, lds_i = 0, lds_gj = 0;
    int inv_threads = 64, mthreads = PZ_MPHASE_THREADS;
    int rad_lds = 1; // collocation m-loop: radial rows staged in LDS
    int n = 0;       // n_r of the class
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
};

struct pz_ensemble {
    std::vector<pz_handle *> runs;
    std::vector<EnsClass> cls;
    Scheme sch;
    int coll = 0;
    int tuned = 0, policy = 0; // autotuned m-loop shape of the large classes
    double policy_ms[4] = {};
    std::vector<void *> allocs;
    std::string err;
    ~pz_ensemble()
    {
        for (EnsClass &c : cls) {
            if (c.gexec) (void)hipGraphExecDestroy(c.gexec);
            if (c.graph) (void)hipGraphDestroy(c.graph);
            if (c.done) (void)hipEventDestroy(c.done);
            if (c.stream) (void)hipStreamDestroy(c.stream);
        }
        for (void *p : allocs) (void)hipFree(p);
    }
    template <class T> T *upload(const std::vector<T> &v)
    {
        void *p = nullptr;
        HIPCHK(hipMalloc(&p, std::max<size_t>(v.size(), 1) * sizeof(T)));
        allocs.push_back(p);
        HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return (T *)p;
    }
};

static void launch_class_step(const pz_ensemble *e, const EnsClass &c)
{
    const Scheme &s = e->sch;
    const PzDev *const *runs = c.druns;
    for (int stage = 1; stage <= s.nstages; stage++) {
        const int exp_calc = s.l_exp_calc[stage - 1];
        if (exp_calc)
            hipLaunchKernelGGL(k_radial_loop_batch, dim3(c.nb_r), dim3(256), c.lds_r, c.stream,
                               PzBatch{runs, c.map_r, (int)c.members.size()}, stage - 1, stage == 1 ? 1 : 0);
        if (stage == 1) {
            hipLaunchKernelGGL(k_courant_batch, dim3(c.nb_c), dim3(256), 0, c.stream, PzBatch{runs, c.map_c, (int)c.members.size()});
            if (e->coll) {
                hipLaunchKernelGGL(k_build_coll_batch, dim3(std::min(c.nb_3, 512)), dim3(1024), c.lds_gj, c.stream,
                                   PzBatch{runs, c.map_3, (int)c.members.size()}, c.nb_3, 0);
            } else {
                hipLaunchKernelGGL(k_build_lhs_batch, dim3(std::min(c.nb_3, 2048)), dim3(64), c.lds_b, c.stream,
                                   PzBatch{runs, c.map_3, (int)c.members.size()}, c.nb_3, 0);
                hipLaunchKernelGGL(k_build_inv_batch, dim3(std::min(c.nb_3, 512)), dim3(c.inv_threads), c.lds_i,
                                   c.stream, PzBatch{runs, c.map_3, (int)c.members.size()}, c.nb_3, 0);
            }
        }
        if (e->coll)
            hipLaunchKernelGGL(c.rad_lds ? k_mcoll_batch<true> : k_mcoll_batch<false>, dim3(c.nb_m), dim3(c.mthreads),
                               c.lds_m, c.stream, PzBatch{runs, c.map_m, (int)c.members.size()}, stage, exp_calc);
        else
            hipLaunchKernelGGL(k_mphase_batch, dim3(c.nb_m), dim3(c.mthreads), c.lds_m, c.stream,
                               PzBatch{runs, c.map_m, (int)c.members.size()}, stage, exp_calc);
    }
}

extern "C" int pz_ensemble_create(pz_handle *const *runs, int nruns, pz_ensemble **out)
{
    if (!runs || nruns < 1 || !out) return -1;
    *out = nullptr;
    pz_ensemble *e = new pz_ensemble();
    try {
        const pz_handle *h0 = runs[0];
        std::vector<int> sizes;
        for (int r = 0; r < nruns; r++) {
            pz_handle *h = runs[r];
            if (!h) throw std::invalid_argument("null run handle");
            if (!h->started) throw std::invalid_argument("every run must be started (pz_start) first");
            if (h->d.fd) throw std::invalid_argument("finite-difference runs are not batched into ensembles");
            if (h->radial_big) throw std::invalid_argument("runs with long azimuthal grids are not batched into ensembles");
            if (h->d.dmat) throw std::invalid_argument("runs with matrix_solve='influence' are not batched into ensembles");
            if (h->sch.l_assembly)
                throw std::invalid_argument("DIRK schemes with an assembly stage are not batched into ensembles");
            if (h->sch.id != h0->sch.id || h->d.coll != h0->d.coll)
                throw std::invalid_argument("the runs of an ensemble must share the time scheme and the radial method");
            e->runs.push_back(h);
            if (std::find(sizes.begin(), sizes.end(), h->N) == sizes.end()) sizes.push_back(h->N);
        }
        std::sort(sizes.begin(), sizes.end());
        e->sch = h0->sch;
        e->coll = h0->d.coll;
        for (int n : sizes) {
            EnsClass c;
            std::vector<const PzDev *> dr;
            std::vector<int2> mr, mc, m3, mm;
            int nw = 16;
            for (int r = 0; r < nruns; r++) {
                const pz_handle *h = runs[r];
                if (h->N != n) continue;
                const int k = (int)dr.size();
                c.members.push_back(r);
                dr.push_back(h->ddev);
                for (int i = 0; i < h->N; i++) mr.push_back(make_int2(k, i));
                mc.push_back(make_int2(k, 0));
                for (int i = 0; i < 3 * h->nm; i++) m3.push_back(make_int2(k, i));
                for (int i = 0; i < h->nm; i++) mm.push_back(make_int2(k, i));
                c.lds_r = std::max(c.lds_r, h->lds_radial);
                c.lds_m = std::max(c.lds_m, h->lds_mphase);
                c.lds_b = std::max(c.lds_b, h->lds_build);
                c.lds_gj = std::max(c.lds_gj, h->lds_gj);
                nw = std::min(nw, h->inv_waves);
            }
            c.lds_i = InvLayout(n, nw).total;
            c.mthreads = mloop_threads(n);
            c.n = n;
            c.inv_threads = 64 * nw;
            c.druns = e->upload(dr);
            c.map_r = e->upload(mr);
            c.map_c = e->upload(mc);
            c.map_3 = e->upload(m3);
            c.map_m = e->upload(mm);
            c.nb_r = (int)mr.size();
            c.nb_c = (int)mc.size();
            c.nb_3 = (int)m3.size();
            c.nb_m = (int)mm.size();
            HIPCHK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&c.done, hipEventDisableTiming));
            e->cls.push_back(c);
        }
    } catch (std::exception &ex) {
        std::fprintf(stderr, "pz_ensemble_create: %s\n", ex.what());
        delete e;
        return -2;
    }
    *out = e;
    return 0;
}

// m-loop shape of the collocation classes with n_r > 65 under a policy:
// 0 the single-run shape (512 threads, radial rows in LDS), 1 256 threads
// (n_r <= 128 only), 2 radial rows and transform tables read through L2, 3
// both (less LDS: more workgroups per CU)
static void ens_policy(pz_ensemble *e, int policy)
{
    int dev = 0, lim = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    for (EnsClass &c : e->cls) {
        const bool big = c.n > 65;
        // the coupled omega-psi product needs 2 n_r <= threads
        c.mthreads = (big && (policy & 1) && 2 * c.n <= 256) ? 256 : mloop_threads(c.n);
        c.rad_lds = (big && (policy & 2)) ? 0 : 1;
        // grids whose full layout does not fit read through L2 under any policy
        if (e->coll && MpLayout(c.n, c.mthreads, true).total > (size_t)lim) c.rad_lds = 0;
        c.lds_m = MpLayout(c.n, c.mthreads, c.rad_lds != 0).total;
        if (c.gexec) { (void)hipGraphExecDestroy(c.gexec); c.gexec = nullptr; }
        if (c.graph) { (void)hipGraphDestroy(c.graph); c.graph = nullptr; }
    }
}

static void ens_capture(pz_ensemble *e)
{
    for (EnsClass &c : e->cls)
        if (!c.gexec) {
            HIPCHK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
            launch_class_step(e, c);
            HIPCHK(hipStreamEndCapture(c.stream, &c.graph));
            HIPCHK(hipGraphInstantiate(&c.gexec, c.graph, nullptr, nullptr, 0));
        }
}

// The autotuner: times a few ensemble steps under each policy from the same
// member states (snapshots of the step state blocks, restored after every
// trial) and keeps the fastest.  The trials may rebuild matrices for
// another dt, so the members' matrices are rebuilt for their restored
// state at the end; the results of the following steps are those of the
// plain step, bit for bit.
static void ens_autotune(pz_ensemble *e)
{
    size_t total = 0;
    for (pz_handle *h : e->runs) total += h->arena_bytes;
    char *snap = nullptr;
    HIPCHK(hipMalloc(&snap, total));
    auto copy = [&](bool save) {
        size_t off = 0;
        for (pz_handle *h : e->runs) {
            HIPCHK(hipMemcpyAsync(save ? snap + off : h->arena, save ? h->arena : snap + off, h->arena_bytes,
                                  hipMemcpyDeviceToDevice, h->stream));
            off += h->arena_bytes;
        }
        for (pz_handle *h : e->runs) HIPCHK(hipStreamSynchronize(h->stream));
    };
    copy(true);
    bool any_big = false;
    for (EnsClass &c : e->cls) any_big |= c.n > 65;
    const int npol = any_big ? 4 : 1, reps = 3;
    double best = 1e300;
    int best_p = 0;
    for (int pol = 0; pol < npol; pol++) {
        ens_policy(e, pol);
        ens_capture(e);
        for (EnsClass &c : e->cls) HIPCHK(hipGraphLaunch(c.gexec, c.stream)); // warm-up (and first builds)
        for (EnsClass &c : e->cls) HIPCHK(hipStreamSynchronize(c.stream));
        const auto t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < reps; i++)
            for (EnsClass &c : e->cls) HIPCHK(hipGraphLaunch(c.gexec, c.stream));
        for (EnsClass &c : e->cls) HIPCHK(hipStreamSynchronize(c.stream));
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / reps;
        e->policy_ms[pol] = ms;
        if (ms < best) { best = ms; best_p = pol; }
        copy(false);
    }
    HIPCHK(hipFree(snap));
    // matrices for the restored states
    for (EnsClass &c : e->cls)
        hipLaunchKernelGGL(k_build_coll_batch, dim3(std::min(c.nb_3, 512)), dim3(1024), c.lds_gj, c.stream,
                           PzBatch{c.druns, c.map_3, (int)c.members.size()}, c.nb_3, 1);
    for (EnsClass &c : e->cls) HIPCHK(hipStreamSynchronize(c.stream));
    HIPCHK(hipGetLastError());
    ens_policy(e, best_p);
    e->policy = best_p;
    e->tuned = 1;
}

extern "C" int pz_ensemble_step(pz_ensemble *e, int nsteps)
{
    if (!e) return -1;
    try {
        // work already queued on the members comes first
        for (pz_handle *h : e->runs) HIPCHK(hipStreamSynchronize(h->stream));
        if (!e->tuned) {
            // PZ_ENSEMBLE_POLICY=k fixes the policy (no tuning)
            const char *fixed = std::getenv("PZ_ENSEMBLE_POLICY");
            if (fixed && e->coll) {
                e->policy = std::atoi(fixed) & 3;
                ens_policy(e, e->policy);
            } else if (e->coll) {
                ens_autotune(e);
            }
            e->tuned = 1;
        }
        for (EnsClass &c : e->cls)
            if (!c.gexec) {
                HIPCHK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
                launch_class_step(e, c);
                HIPCHK(hipStreamEndCapture(c.stream, &c.graph));
                HIPCHK(hipGraphInstantiate(&c.gexec, c.graph, nullptr, nullptr, 0));
            }
        // the classes advance side by side, each on its own stream
        for (int i = 0; i < nsteps; i++)
            for (EnsClass &c : e->cls) HIPCHK(hipGraphLaunch(c.gexec, c.stream));
        HIPCHK(hipGetLastError());
        // later calls on a member wait for its class's work
        for (EnsClass &c : e->cls) {
            HIPCHK(hipEventRecord(c.done, c.stream));
            for (int r : c.members) HIPCHK(hipStreamWaitEvent(e->runs[r]->stream, c.done, 0));
        }
    } catch (std::exception &ex) {
        e->err = ex.what();
        return -4;
    }
    return 0;
}

extern "C" int pz_ensemble_destroy(pz_ensemble *e)
{
    if (e) {
        for (EnsClass &c : e->cls) (void)hipStreamSynchronize(c.stream);
        delete e;
    }
    return 0;
}

extern "C" const char *pz_ensemble_last_error(const pz_ensemble *e) { return e ? e->err.c_str() : "null ensemble"; }

extern "C" int pz_ensemble_policy(const pz_ensemble *e, int *policy, double *trial_ms)
{
    if (!e || !policy) return -1;
    *policy = e->policy;
    if (trial_ms)
        for (int i = 0; i < 4; i++) trial_ms[i] = e->policy_ms[i];
    return e->tuned ? 0 : 1;
}
