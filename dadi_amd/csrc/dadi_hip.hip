// dadi_hip.hip -- MI355X (gfx950) kernels and C ABI for dadi's ADI time
// stepping of the allele-frequency density phi (1-5 populations) and the
// phi -> spectrum sampling integral.  See include/dadi_hip.h for the ABI.
//
// Design (CDNA4, FP64 VALU, no MFMA):
//  * One lane solves one tridiagonal line with the Numerical-Recipes Thomas
//    recurrence (dadi/tridiag.c:14-34).  Lines of the strided axes are taken
//    by consecutive lanes at consecutive addresses, so every row load/store of
//    a wave is one coalesced 512-byte access.
//  * Coefficients are never materialised: the per-axis 1D vectors
//    (dfactor, xInt, V/(2dx) at both ends of each interval, selection term)
//    live in LDS, are identical for every lane (broadcast reads), and each
//    lane adds only its own migration term from the other axes' coordinates
//    (dadi/integration_shared.c:4-93 folded into the forward sweep).
//  * The forward elimination writes u' over phi in place and keeps the
//    gamma_j column of its lane in LDS (stride 64 doubles: conflict free).
//  * The whole time loop of an instance runs on the device.  The nD kernel is
//    persistent: each workgroup pulls whole instances from a device-side
//    queue; the 1D kernel does the same per lane.
//  * Compiled with -ffp-contract=off and IEEE division, in the reference's
//    operation order: results are bit-identical to the reference C.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "dadi_hip.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define DH_HIP(expr)                                                            \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess)                                                   \
            return fail(DH_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int g_num_cu = 0;
long long g_wall_khz = 0;   // rate of wall_clock64()
size_t g_max_lds = 0;
bool g_inited = false;

constexpr int LDS_TABLES_PER_AXIS = 3;  // W0, W1, S
constexpr int LDS_GRID_TABLES = 4;      // x, xInt, dfactor, dx

}  // namespace

// ---------------------------------------------------------------------------
// Device arithmetic, each in the reference's association order.
// ---------------------------------------------------------------------------

// V(x) = 1/nu * x * (1-x), dadi/integration_shared.c:4-6
__device__ __forceinline__ double d_vfunc(double x, double nu)
{
    return 1. / nu * x * (1. - x);
}

// selection part of M, dadi/integration_shared.c:12-30; for an X chromosome
// (chrom = 1) gamma * 4/3 * (0.5 + h + x (1-2h)) x (1-x),
// dadi/Integration.py:1246-1247
__device__ __forceinline__ double d_sel(double x, double gamma, double h, int chrom = 0)
{
    if (chrom) return gamma * 4. / 3. * (0.5 + h + x * (1. - 2 * h)) * x * (1. - x);
    return gamma * 2 * (h + (1. - 2 * h) * x) * x * (1. - x);
}

// M = m_0 (c_0 - x) + m_1 (c_1 - x) + ... + sel, summed left to right
template <int NO>
__device__ __forceinline__ double d_M(const double *c, const double *m, double x, double sel)
{
    if constexpr (NO == 0) {
        return sel;
    } else {
        double acc = m[0] * (c[0] - x);
#pragma unroll
        for (int q = 1; q < NO; ++q) acc = acc + m[q] * (c[q] - x);
        return acc + sel;
    }
}

// Device pointers stored in instance records are read as global-address-space
// pointers, so the compiler emits global (not flat) memory instructions for
// phi, the grid and the schedule: their completion is counted apart from LDS
// traffic, and LDS waits no longer wait for phi rows still in flight.
template <class T>
__device__ __forceinline__ T *gbl(T *const &field)
{
    using GP = __attribute__((address_space(1))) T *;
    return (T *)*(const GP *)&field;
}

// Chang and Cooper's delta of one interval (use_delj_trick): wj = 2 M dx,
// epsj = exp(wj/VInt), delta = (-epsj wj + epsj VInt - VInt)/(wj - epsj wj).
// The C sweeps (time-varying order) fall back to 1/2 when epsj == 1 or
// wj == 0 (dadi/integration_shared.c:52-70); the numpy coefficient arrays of
// the constant-parameter paths replace NaN and inf by 1/2
// (dadi/Integration.py:960-984).  exp is the device's, so results agree
// with the reference to rounding, not bit for bit.
template <int ORDER>
__device__ __forceinline__ double d_delj(double M, double VI, double dx)
{
    const double wj = 2 * M * dx;
    const double epsj = exp(wj / VI);
    if constexpr (ORDER == DH_ORDER_TV) {
        if ((epsj != 1.0) && (wj != 0)) return (-epsj * wj + epsj * VI - VI) / (wj - epsj * wj);
        return 0.5;
    } else {
        const double d = (-epsj * wj + epsj * VI - VI) / (wj - epsj * wj);
        return (__builtin_isnan(d) || __builtin_isinf(d)) ? 0.5 : d;
    }
}

// Per-interval tables of the delta trick: V at the interval midpoints and
// the interval widths, with the stride of the other coefficient tables.
struct DeljTabs {
    const double *VI = nullptr;
    const double *DX = nullptr;
};

// Compile-time tunables (defaults measured best on MI355X):
//  DH_PF       rows fetched ahead of use in the Thomas sweeps.  The store of
//              row j and the load of row j+PF go to different addresses, so
//              issuing the loads a block early hides their latency without
//              changing any result.
//  DH_STAGE_C  columns per staging chunk of the contiguous-axis tiles (8
//              doubles = 64 B: fewer registers in flight than a 128-B line).
//  DH_WPE      minimum waves per SIMD the nD kernel is compiled for (bounds
//              its registers: 3 -> at most 168 VGPRs).
#ifndef DH_PF
#define DH_PF 4
#endif
#ifndef DH_STAGE_C
#define DH_STAGE_C 8
#endif
#ifndef DH_WPE
#define DH_WPE 3
#endif
constexpr int PF = DH_PF;
// Rows per block of the multi-workgroup sweeps (k_sweep_axis, k_sweep_batch):
// one wave per SIMD there, so nothing hides a row's latency but the rows of
// a block interleaving (3-population sweeps on 40-64 points, MI355X: 8 rows
// 11-14% faster per step than 4)
#ifndef DH_SWEEP_PF
#define DH_SWEEP_PF 8
#endif
constexpr int SWEEP_PF = DH_SWEEP_PF;
constexpr int STAGE_C = DH_STAGE_C;

// Coefficients a_j, b_j and c_{j-1} of row j >= 1 (compute_abc_nobc,
// dadi/integration_shared.c:72-93, with the boundary term of the last row);
// pa/pc carry atemp/ctemp of interface j-1 and are advanced to interface j.
template <int NO, int ORDER, int DJ = 0>
__device__ __forceinline__ void row_coeffs(int j, int n, const double *__restrict__ xInt,
                                           const double *__restrict__ df,
                                           const double *__restrict__ W0,
                                           const double *__restrict__ W1,
                                           const double *__restrict__ S, long long ts,
                                           const double *c, const double *m, double inv_dt,
                                           bool bcl_on, double bcl, double &pa, double &pc,
                                           double &a, double &b, double &cm, DeljTabs dj = {})
{
    const double dfj = df[j * ts];
    a = -dfj * pa;
    cm = -df[(j - 1) * ts] * pc;
    double na = 0., nc = 0.;
    if (j < n - 1) {
        const double Mj = d_M<NO>(c, m, xInt[j * ts], S[j * ts]);
        const double dl = DJ ? d_delj<ORDER>(Mj, dj.VI[j * ts], dj.DX[j * ts]) : 0.5;
        na = Mj * dl + W0[j * ts];
        nc = -Mj * (1 - dl) + W1[j * ts];
        if constexpr (ORDER == DH_ORDER_TV)
            b = (inv_dt + dfj * pc) + dfj * na;
        else
            b = (dfj * na + dfj * pc) + inv_dt;
    } else {
        if constexpr (ORDER == DH_ORDER_TV) {
            b = inv_dt + dfj * pc;
            if (bcl_on) b += bcl;
        } else {
            b = dfj * pc;
            if (bcl_on) b += bcl;
            b = b + inv_dt;
        }
    }
    pa = na;
    pc = nc;
}

// Row 0: coefficients, boundary term and the first Thomas step.
template <int NO, int ORDER, int DJ = 0>
__device__ __forceinline__ void row0(double p0, const double *__restrict__ xInt,
                                     const double *__restrict__ df,
                                     const double *__restrict__ W0,
                                     const double *__restrict__ W1,
                                     const double *__restrict__ S, const double *c,
                                     const double *m, double dt, double inv_dt, bool bcf_on,
                                     double bcf, double &bet, double &u, double &pa, double &pc,
                                     DeljTabs dj = {})
{
    const double M = d_M<NO>(c, m, xInt[0], S[0]);
    const double dl = DJ ? d_delj<ORDER>(M, dj.VI[0], dj.DX[0]) : 0.5;
    pa = M * dl + W0[0];
    pc = -M * (1 - dl) + W1[0];
    double b;
    if constexpr (ORDER == DH_ORDER_TV) {
        b = inv_dt + df[0] * pa;
        if (bcf_on) b += bcf;
    } else {
        b = df[0] * pa;
        if (bcf_on) b += bcf;
        b = b + inv_dt;
    }
    const double r = (ORDER == DH_ORDER_PRE) ? inv_dt * p0 : p0 / dt;
    bet = b;
    u = r / bet;
}

// One forward-elimination row j >= 1 (dadi/tridiag.c:25-29); pj is phi at
// row j.  Returns gamma_j; updates bet, u (= u'_j) and the coefficient carry.
template <int NO, int ORDER, int DJ = 0>
__device__ __forceinline__ double fwd_core(int j, int n, double pj,
                                           const double *__restrict__ xInt,
                                           const double *__restrict__ df,
                                           const double *__restrict__ W0,
                                           const double *__restrict__ W1,
                                           const double *__restrict__ S, long long ts,
                                           const double *c, const double *m, double dt,
                                           double inv_dt, bool bcl_on, double bcl,
                                           double &bet, double &u, double &pa, double &pc,
                                           DeljTabs dj = {})
{
    double a, b, cm;
    row_coeffs<NO, ORDER, DJ>(j, n, xInt, df, W0, W1, S, ts, c, m, inv_dt, bcl_on, bcl, pa, pc,
                              a, b, cm, dj);
    const double g = cm / bet;
    bet = b - a * g;
    const double r = (ORDER == DH_ORDER_PRE) ? inv_dt * pj : pj / dt;
    u = (r - a * u) / bet;
    return g;
}

template <int NO, int ORDER, int DJ = 0>
__device__ __forceinline__ void fwd_row(int j, int n, double pj,
                                        const double *__restrict__ xInt,
                                        const double *__restrict__ df,
                                        const double *__restrict__ W0,
                                        const double *__restrict__ W1,
                                        const double *__restrict__ S, long long ts,
                                        const double *c, const double *m, double dt,
                                        double inv_dt, bool bcl_on, double bcl,
                                        double &bet, double &u, double &pa, double &pc,
                                        double *__restrict__ p, long long s,
                                        double *__restrict__ gam, long long gs,
                                        DeljTabs dj = {})
{
    const double g = fwd_core<NO, ORDER, DJ>(j, n, pj, xInt, df, W0, W1, S, ts, c, m, dt,
                                             inv_dt, bcl_on, bcl, bet, u, pa, pc, dj);
    gam[j * gs] = g;
    p[j * s] = u;
}

// Gammas of indices cb+1 .. cb+CK recomputed from the checkpoint bet_cb with
// the forward sweep's own operations (bit-identical values).
template <int NO, int ORDER, int CK>
__device__ __forceinline__ void regamma(int cb, int n, double bt,
                                        const double *__restrict__ xInt,
                                        const double *__restrict__ df,
                                        const double *__restrict__ W0,
                                        const double *__restrict__ W1,
                                        const double *__restrict__ S, long long ts,
                                        const double *c, const double *m, double inv_dt,
                                        bool bcl_on, double bcl, double (&g)[CK])
{
    const double Mc = d_M<NO>(c, m, xInt[cb * ts], S[cb * ts]);
    double qa = Mc * 0.5 + W0[cb * ts];
    double qc = -Mc * (1 - 0.5) + W1[cb * ts];
#pragma unroll
    for (int q = 0; q < CK; ++q) {
        const int j = cb + 1 + q;
        g[q] = 0.;
        if (j <= n - 1) {
            double a, b, cm;
            row_coeffs<NO, ORDER>(j, n, xInt, df, W0, W1, S, ts, c, m, inv_dt, bcl_on, bcl, qa,
                                  qc, a, b, cm);
            g[q] = cm / bt;
            bt = b - a * g[q];
        }
    }
}

// BP interior rows j0 .. j0+BP-1 (all below the last row) of the forward
// elimination: fwd_row's operations, with the block's table entries loaded
// before any of its stores (the gamma column shares the LDS with the tables,
// so loads after a store could not move above it: each row waited for its
// own table reads).  Used by the multi-workgroup sweeps (one wave per SIMD);
// the persistent kernels, register-bound, keep fwd_row.
template <int NO, int ORDER, int BP>
__device__ __forceinline__ void fwd_block_interior(
    int j0, const double (&cur)[BP], const double *__restrict__ xInt,
    const double *__restrict__ df, const double *__restrict__ W0,
    const double *__restrict__ W1, const double *__restrict__ S, long long ts,
    const double *c, const double *m, double dt, double inv_dt, double &bet, double &u,
    double &pa, double &pc, double *__restrict__ p, long long s, double *__restrict__ gam,
    long long gs)
{
    double tdf[BP + 1], txi[BP], tsl[BP], tw0[BP], tw1[BP];
    tdf[0] = df[(j0 - 1) * ts];
#pragma unroll
    for (int q = 0; q < BP; ++q) {
        const long long o = (long long)(j0 + q) * ts;
        tdf[q + 1] = df[o];
        txi[q] = xInt[o];
        tsl[q] = S[o];
        tw0[q] = W0[o];
        tw1[q] = W1[o];
    }
#pragma unroll
    for (int q = 0; q < BP; ++q) {
        // row_coeffs (j < n - 1, no delta trick) and fwd_core
        const double dfj = tdf[q + 1];
        const double a = -dfj * pa;
        const double cm = -tdf[q] * pc;
        const double Mj = d_M<NO>(c, m, txi[q], tsl[q]);
        const double na = Mj * 0.5 + tw0[q];
        const double nc = -Mj * (1 - 0.5) + tw1[q];
        double b;
        if constexpr (ORDER == DH_ORDER_TV)
            b = (inv_dt + dfj * pc) + dfj * na;
        else
            b = (dfj * na + dfj * pc) + inv_dt;
        pa = na;
        pc = nc;
        const double g = cm / bet;
        bet = b - a * g;
        const double r = (ORDER == DH_ORDER_PRE) ? inv_dt * cur[q] : cur[q] / dt;
        u = (r - a * u) / bet;
        gam[(j0 + q) * gs] = g;
        p[(j0 + q) * s] = u;
    }
}

// One line: inject (optional), build a/b/c on the fly, Thomas forward with u'
// over phi in place and gamma_j in `gam`, then back substitution.
//   ts: stride of the 1D tables; gs: stride of the gamma column.
template <int NO, int ORDER, int DJ = 0, int BP = PF, bool IB = false>
__device__ __forceinline__ void solve_line(
    double *__restrict__ p, long long s, int n,
    const double *__restrict__ xInt, const double *__restrict__ df,
    const double *__restrict__ W0, const double *__restrict__ W1,
    const double *__restrict__ S, long long ts,
    const double *c, const double *m,
    double dt, double inv_dt,
    bool bcf_on, double bcf, bool bcl_on, double bcl,
    int inj_row, double inj,
    double *__restrict__ gam, long long gs, DeljTabs dj = {})
{
    if (inj_row >= 0) p[inj_row * s] = p[inj_row * s] + inj;

    // rows 1..BP in flight while row 0 is formed
    double nxt[BP];
#pragma unroll
    for (int q = 0; q < BP; ++q) {
        const int jj = 1 + q;
        nxt[q] = p[(jj < n ? jj : n - 1) * s];
    }

    // row 0 (compute_abc_nobc, dadi/integration_shared.c:72-93)
    double bet, u, pa, pc;
    row0<NO, ORDER, DJ>(p[0], xInt, df, W0, W1, S, c, m, dt, inv_dt, bcf_on, bcf, bet, u, pa,
                        pc, dj);
    p[0] = u;

    // forward elimination, BP rows per block, the next block's rows loading.
    // A whole block has no per-row branch, so one basic block holds its rows
    // and the compiler interleaves row j's u division with row j+1's
    // coefficients and gamma division (waves issue in order: with a branch
    // per row each row's two divisions ran back to back)
    for (int j0 = 1; j0 < n; j0 += BP) {
        double cur[BP];
#pragma unroll
        for (int q = 0; q < BP; ++q) cur[q] = nxt[q];
#pragma unroll
        for (int q = 0; q < BP; ++q) {
            const int jj = j0 + BP + q;
            nxt[q] = p[(jj < n ? jj : n - 1) * s];
        }
        if (IB && !DJ && j0 + BP <= n - 1) {
            fwd_block_interior<NO, ORDER, BP>(j0, cur, xInt, df, W0, W1, S, ts, c, m, dt,
                                              inv_dt, bet, u, pa, pc, p, s, gam, gs);
        } else if (j0 + BP <= n) {
#pragma unroll
            for (int q = 0; q < BP; ++q)
                fwd_row<NO, ORDER, DJ>(j0 + q, n, cur[q], xInt, df, W0, W1, S, ts, c, m, dt,
                                       inv_dt, bcl_on, bcl, bet, u, pa, pc, p, s, gam, gs, dj);
        } else {
#pragma unroll
            for (int q = 0; q < BP; ++q) {
                const int j = j0 + q;
                if (j < n)
                    fwd_row<NO, ORDER, DJ>(j, n, cur[q], xInt, df, W0, W1, S, ts, c, m, dt,
                                           inv_dt, bcl_on, bcl, bet, u, pa, pc, p, s, gam, gs,
                                           dj);
            }
        }
    }

    // back substitution (dadi/tridiag.c:31-33), rows and gammas loaded a
    // block ahead
    double bp[BP], bg[BP];
#pragma unroll
    for (int q = 0; q < BP; ++q) {
        const int jj = n - 2 - q;
        const int jc = jj >= 0 ? jj : 0;
        bp[q] = p[jc * s];
        bg[q] = gam[(jc + 1) * gs];
    }
    for (int j0 = n - 2; j0 >= 0; j0 -= BP) {
        double cp[BP], cg[BP];
#pragma unroll
        for (int q = 0; q < BP; ++q) {
            cp[q] = bp[q];
            cg[q] = bg[q];
        }
#pragma unroll
        for (int q = 0; q < BP; ++q) {
            const int jj = j0 - BP - q;
            const int jc = jj >= 0 ? jj : 0;
            bp[q] = p[jc * s];
            bg[q] = gam[(jc + 1) * gs];
        }
        if (j0 + 1 >= BP) {
#pragma unroll
            for (int q = 0; q < BP; ++q) {
                const double v = cp[q] - cg[q] * u;
                p[(j0 - q) * s] = v;
                u = v;
            }
        } else {
#pragma unroll
            for (int q = 0; q < BP; ++q) {
                const int j = j0 - q;
                if (j >= 0) {
                    const double v = cp[q] - cg[q] * u;
                    p[j * s] = v;
                    u = v;
                }
            }
        }
    }
}

// Thomas with checkpoints: the forward sweep keeps only bet at rows
// 0, CK, 2CK, ... (not the whole gamma column); the back substitution
// recomputes each chunk's gammas from its checkpoint with the very same
// operations, holding them in registers.  Same bits as solve_line, ~CK times
// less LDS per lane, more VALU.
template <int NO, int ORDER, int CK>
__device__ __forceinline__ void solve_line_ck(
    double *__restrict__ p, long long s, int n,
    const double *__restrict__ xInt, const double *__restrict__ df,
    const double *__restrict__ W0, const double *__restrict__ W1,
    const double *__restrict__ S, long long ts,
    const double *c, const double *m,
    double dt, double inv_dt,
    bool bcf_on, double bcf, bool bcl_on, double bcl,
    int inj_row, double inj,
    double *__restrict__ chk, long long gs)
{
    if (inj_row >= 0) p[inj_row * s] = p[inj_row * s] + inj;
    double nxt[PF];
#pragma unroll
    for (int q = 0; q < PF; ++q) {
        const int jj = 1 + q;
        nxt[q] = p[(jj < n ? jj : n - 1) * s];
    }
    double bet, u, pa, pc;
    row0<NO, ORDER>(p[0], xInt, df, W0, W1, S, c, m, dt, inv_dt, bcf_on, bcf, bet, u, pa, pc);
    p[0] = u;
    chk[0] = bet;
    for (int j0 = 1; j0 < n; j0 += PF) {
        double cur[PF];
#pragma unroll
        for (int q = 0; q < PF; ++q) cur[q] = nxt[q];
#pragma unroll
        for (int q = 0; q < PF; ++q) {
            const int jj = j0 + PF + q;
            nxt[q] = p[(jj < n ? jj : n - 1) * s];
        }
#pragma unroll
        for (int q = 0; q < PF; ++q) {
            const int j = j0 + q;
            if (j < n) {
                fwd_core<NO, ORDER>(j, n, cur[q], xInt, df, W0, W1, S, ts, c, m, dt, inv_dt,
                                    bcl_on, bcl, bet, u, pa, pc);
                p[j * s] = u;
                if (j % CK == 0) chk[(j / CK) * gs] = bet;
            }
        }
    }
    // back substitution, one chunk of gamma indices [cb+1, cb+CK] at a time
    for (int cb = ((n - 2) / CK) * CK; cb >= 0; cb -= CK) {
        double row[CK];
#pragma unroll
        for (int q = 0; q < CK; ++q) {
            const int jr = cb + q;
            row[q] = p[(jr <= n - 2 ? jr : n - 2) * s];
        }
        double g[CK];
        regamma<NO, ORDER, CK>(cb, n, chk[(cb / CK) * gs], xInt, df, W0, W1, S, ts, c, m,
                               inv_dt, bcl_on, bcl, g);
#pragma unroll
        for (int q = CK - 1; q >= 0; --q) {
            const int jr = cb + q;
            if (jr <= n - 2) {
                const double v = row[q] - g[q] * u;
                p[jr * s] = v;
                u = v;
            }
        }
    }
}

// Orders loads/stores to LDS between lanes of one wave: a compiler barrier
// (LDS operations of a wave complete in issue order).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// 64 lines of the contiguous axis solved by one wave: the lines are 64
// consecutive rows of n doubles, moved between HBM and LDS in chunks of C
// columns with every lane touching consecutive addresses (each row chunk is
// one 128-byte run), while each lane runs its own Thomas row out of LDS.
// Checkpointed gammas (CK); u' written back between the two passes except for
// the last staging chunk, which stays in LDS.
// Row addressing of a staged contiguous-axis tile.  Element q of a lane in a
// staging chunk is row q*(64/C) + lane/C, column lane%C of the chunk.
//  TileRows:   64 consecutive lines of one phi starting at `tile`.
//  PackedRows: row groups of 64/C lines, each group inside one instance
//              (base[q] = its first row; valid[q] = rows of it in use).
template <int C>
struct TileRows {
    double *tile;
    int rows_active;
    long long ld;   // row pitch
    __device__ __forceinline__ bool ok(int q, int lane, int cw) const
    {
        const int row = q * (64 / C) + lane / C, col = lane % C;
        return row < rows_active && col < cw;
    }
    __device__ __forceinline__ double *at(int q, int lane, int j0, bool ok) const
    {
        const int row = q * (64 / C) + lane / C, col = lane % C;
        return tile + (ok ? (long long)row * ld + j0 + col : 0);
    }
};

template <int C>
struct PackedRows {
    double *base[C];
    int valid[C];
    long long ld;   // row pitch
    __device__ __forceinline__ bool ok(int q, int lane, int cw) const
    {
        return lane / C < valid[q] && lane % C < cw;
    }
    __device__ __forceinline__ double *at(int q, int lane, int j0, bool ok) const
    {
        return base[q] + (ok ? (long long)(lane / C) * ld + j0 + lane % C : 0);
    }
};

template <int NO, int ORDER, int CK, int C, class Rows>
__device__ __forceinline__ void solve_tile_staged(
    const Rows &R, int n,
    const double *__restrict__ xInt, const double *__restrict__ df,
    const double *__restrict__ W0, const double *__restrict__ W1,
    const double *__restrict__ S,
    const double *c, const double *m, double dt, double inv_dt,
    bool bcf_on, double bcf, bool bcl_on, double bcl, int inj_row, double inj,
    double *__restrict__ chk, double *__restrict__ buf, int lane)
{
    static_assert(C % CK == 0, "staging chunk must hold whole checkpoint chunks");
    constexpr int LD = C + 1;
    double *mine = buf + lane * LD;
    const int nch = (n + C - 1) / C;
    double bet = 0., u = 0., pa = 0., pc = 0.;
    for (int k = 0; k < nch; ++k) {
        const int j0 = k * C;
        const int cw = n - j0 < C ? n - j0 : C;
        {
            double v[C];
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const bool ok = R.ok(q, lane, cw);
                v[q] = *R.at(q, lane, j0, ok);
                v[q] = ok ? v[q] : 0.;
            }
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const int e = q * 64 + lane, row = e / C, col = e - row * C;
                buf[row * LD + col] = v[q];
            }
        }
        wave_sync();
        if (k == 0 && inj_row >= 0) mine[inj_row] = mine[inj_row] + inj;
        for (int cc = 0; cc < cw; ++cc) {
            const int j = j0 + cc;
            const double pj = mine[cc];
            if (j == 0) {
                row0<NO, ORDER>(pj, xInt, df, W0, W1, S, c, m, dt, inv_dt, bcf_on, bcf, bet, u,
                                pa, pc);
                chk[0] = bet;
            } else {
                fwd_core<NO, ORDER>(j, n, pj, xInt, df, W0, W1, S, 1, c, m, dt, inv_dt, bcl_on,
                                    bcl, bet, u, pa, pc);
                if (j % CK == 0) chk[(j / CK) * 64] = bet;
            }
            mine[cc] = u;
        }
        wave_sync();
        if (k == nch - 1) break;   // the last chunk is consumed from LDS
        {
            double v[C];
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const int e = q * 64 + lane, row = e / C, col = e - row * C;
                v[q] = buf[row * LD + col];
            }
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const bool ok = R.ok(q, lane, cw);
                if (ok) *R.at(q, lane, j0, ok) = v[q];
            }
        }
        wave_sync();
    }
    for (int k = nch - 1; k >= 0; --k) {
        const int j0 = k * C;
        const int cw = n - j0 < C ? n - j0 : C;
        if (k < nch - 1) {
            double v[C];
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const bool ok = R.ok(q, lane, cw);
                v[q] = *R.at(q, lane, j0, ok);
                v[q] = ok ? v[q] : 0.;
            }
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const int e = q * 64 + lane, row = e / C, col = e - row * C;
                buf[row * LD + col] = v[q];
            }
            wave_sync();
        }
        const int top = (n - 2 < j0 + cw - 1) ? n - 2 : j0 + cw - 1;
        for (int cb = top >= j0 ? (top / CK) * CK : j0 - CK; cb >= j0; cb -= CK) {
            double g[CK];
            regamma<NO, ORDER, CK>(cb, n, chk[(cb / CK) * 64], xInt, df, W0, W1, S, 1, c, m,
                                   inv_dt, bcl_on, bcl, g);
#pragma unroll
            for (int q = CK - 1; q >= 0; --q) {
                const int jr = cb + q;
                if (jr <= n - 2) {
                    const double v = mine[jr - j0] - g[q] * u;
                    mine[jr - j0] = v;
                    u = v;
                }
            }
        }
        wave_sync();
        {
            double v[C];
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const int e = q * 64 + lane, row = e / C, col = e - row * C;
                v[q] = buf[row * LD + col];
            }
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const bool ok = R.ok(q, lane, cw);
                if (ok) *R.at(q, lane, j0, ok) = v[q];
            }
        }
        wave_sync();
    }
}

// Contiguous-axis tile with each lane's gamma column in a global workspace
// (gam: this lane's column, stride 64 doubles between rows): the back
// substitution reads the forward sweep's gammas instead of recomputing them.
// The last staging chunk stays in LDS between the two passes.
template <int NO, int ORDER, int C>
__device__ __forceinline__ void solve_tile_staged_gg(
    double *__restrict__ tile, int rows_active, int n, long long ld,
    const double *__restrict__ xInt, const double *__restrict__ df,
    const double *__restrict__ W0, const double *__restrict__ W1,
    const double *__restrict__ S,
    const double *c, const double *m, double dt, double inv_dt,
    bool bcf_on, double bcf, bool bcl_on, double bcl, int inj_row, double inj,
    double *__restrict__ gam, double *__restrict__ buf, int lane)
{
    constexpr int LD = C + 1;
    double *mine = buf + lane * LD;
    const int nch = (n + C - 1) / C;
    double bet = 0., u = 0., pa = 0., pc = 0.;
    for (int k = 0; k < nch; ++k) {
        const int j0 = k * C;
        const int cw = n - j0 < C ? n - j0 : C;
        {
            double v[C];
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const int e = q * 64 + lane, row = e / C, col = e - row * C;
                const bool ok = row < rows_active && col < cw;
                v[q] = tile[ok ? (long long)row * ld + j0 + col : 0];
                v[q] = ok ? v[q] : 0.;
            }
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const int e = q * 64 + lane, row = e / C, col = e - row * C;
                buf[row * LD + col] = v[q];
            }
        }
        wave_sync();
        if (k == 0 && inj_row >= 0) mine[inj_row] = mine[inj_row] + inj;
        for (int cc = 0; cc < cw; ++cc) {
            const int j = j0 + cc;
            const double pj = mine[cc];
            if (j == 0) {
                row0<NO, ORDER>(pj, xInt, df, W0, W1, S, c, m, dt, inv_dt, bcf_on, bcf, bet, u,
                                pa, pc);
            } else {
                gam[j * 64] = fwd_core<NO, ORDER>(j, n, pj, xInt, df, W0, W1, S, 1, c, m, dt,
                                                  inv_dt, bcl_on, bcl, bet, u, pa, pc);
            }
            mine[cc] = u;
        }
        wave_sync();
        if (k == nch - 1) break;   // the last chunk is consumed from LDS
        {
            double v[C];
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const int e = q * 64 + lane, row = e / C, col = e - row * C;
                v[q] = buf[row * LD + col];
            }
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const int e = q * 64 + lane, row = e / C, col = e - row * C;
                if (row < rows_active && col < cw) tile[(long long)row * ld + j0 + col] = v[q];
            }
        }
        wave_sync();
    }
    for (int k = nch - 1; k >= 0; --k) {
        const int j0 = k * C;
        const int cw = n - j0 < C ? n - j0 : C;
        double g[C];
#pragma unroll
        for (int q = 0; q < C; ++q) {
            const int jj = j0 + q + 1;
            g[q] = gam[(jj <= n - 1 ? jj : n - 1) * 64];
        }
        if (k < nch - 1) {
            double v[C];
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const int e = q * 64 + lane, row = e / C, col = e - row * C;
                const bool ok = row < rows_active && col < cw;
                v[q] = tile[ok ? (long long)row * ld + j0 + col : 0];
                v[q] = ok ? v[q] : 0.;
            }
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const int e = q * 64 + lane, row = e / C, col = e - row * C;
                buf[row * LD + col] = v[q];
            }
            wave_sync();
        }
#pragma unroll
        for (int q = C - 1; q >= 0; --q) {
            const int jr = j0 + q;
            if (q < cw && jr <= n - 2) {
                const double v = mine[q] - g[q] * u;
                mine[q] = v;
                u = v;
            }
        }
        wave_sync();
        {
            double v[C];
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const int e = q * 64 + lane, row = e / C, col = e - row * C;
                v[q] = buf[row * LD + col];
            }
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const int e = q * 64 + lane, row = e / C, col = e - row * C;
                if (row < rows_active && col < cw) tile[(long long)row * ld + j0 + col] = v[q];
            }
        }
        wave_sync();
    }
}

// Recompute variant: the forward sweep keeps (bet, u') only at rows
// 0, CK, 2CK, ... and writes nothing to phi but its last row; the back
// substitution reloads each chunk's original phi rows and recomputes the
// chunk's gammas and u' from the checkpoint with the forward sweep's own
// operations (same bits), so phi makes one read-read-write trip per sweep
// instead of read-write-read-write.
template <int NO, int ORDER, int CK>
__device__ __forceinline__ void regamma_u(int cb, int n, double bt, double u0,
                                          const double (&ph)[CK],
                                          const double *__restrict__ xInt,
                                          const double *__restrict__ df,
                                          const double *__restrict__ W0,
                                          const double *__restrict__ W1,
                                          const double *__restrict__ S, long long ts,
                                          const double *c, const double *m, double dt,
                                          double inv_dt, bool bcl_on, double bcl,
                                          double (&g)[CK], double (&up)[CK])
{
    const double Mc = d_M<NO>(c, m, xInt[cb * ts], S[cb * ts]);
    double qa = Mc * 0.5 + W0[cb * ts];
    double qc = -Mc * (1 - 0.5) + W1[cb * ts];
    double uu = u0;
    up[0] = u0;
#pragma unroll
    for (int q = 0; q < CK; ++q) {
        const int j = cb + 1 + q;
        g[q] = 0.;
        if (q + 1 < CK) up[q + 1] = 0.;
        if (j <= n - 1) {
            double a, b, cm;
            row_coeffs<NO, ORDER>(j, n, xInt, df, W0, W1, S, ts, c, m, inv_dt, bcl_on, bcl, qa,
                                  qc, a, b, cm);
            g[q] = cm / bt;
            bt = b - a * g[q];
            if (q + 1 < CK && j <= n - 2) {
                const double r = (ORDER == DH_ORDER_PRE) ? inv_dt * ph[q + 1] : ph[q + 1] / dt;
                uu = (r - a * uu) / bt;
                up[q + 1] = uu;
            }
        }
    }
}

template <int NO, int ORDER, int CK>
__device__ __forceinline__ void solve_line_ru(
    double *__restrict__ p, long long s, int n,
    const double *__restrict__ xInt, const double *__restrict__ df,
    const double *__restrict__ W0, const double *__restrict__ W1,
    const double *__restrict__ S, long long ts,
    const double *c, const double *m,
    double dt, double inv_dt,
    bool bcf_on, double bcf, bool bcl_on, double bcl,
    int inj_row, double inj,
    double *__restrict__ chk, long long gs, int col)
{
    double *chku = chk + (long long)col * gs;
    if (inj_row >= 0) p[inj_row * s] = p[inj_row * s] + inj;
    double nxt[PF];
#pragma unroll
    for (int q = 0; q < PF; ++q) {
        const int jj = 1 + q;
        nxt[q] = p[(jj < n ? jj : n - 1) * s];
    }
    double bet, u, pa, pc;
    row0<NO, ORDER>(p[0], xInt, df, W0, W1, S, c, m, dt, inv_dt, bcf_on, bcf, bet, u, pa, pc);
    chk[0] = bet;
    chku[0] = u;
    for (int j0 = 1; j0 < n; j0 += PF) {
        double cur[PF];
#pragma unroll
        for (int q = 0; q < PF; ++q) cur[q] = nxt[q];
#pragma unroll
        for (int q = 0; q < PF; ++q) {
            const int jj = j0 + PF + q;
            nxt[q] = p[(jj < n ? jj : n - 1) * s];
        }
#pragma unroll
        for (int q = 0; q < PF; ++q) {
            const int j = j0 + q;
            if (j < n) {
                fwd_core<NO, ORDER>(j, n, cur[q], xInt, df, W0, W1, S, ts, c, m, dt, inv_dt,
                                    bcl_on, bcl, bet, u, pa, pc);
                if (j % CK == 0) {
                    chk[(j / CK) * gs] = bet;
                    chku[(j / CK) * gs] = u;
                }
            }
        }
    }
    p[(n - 1) * s] = u;   // u'_{n-1} is already the solution
    for (int cb = ((n - 2) / CK) * CK; cb >= 0; cb -= CK) {
        double ph[CK];
#pragma unroll
        for (int q = 0; q < CK; ++q) {
            const int jr = cb + q;
            ph[q] = p[(jr <= n - 2 ? jr : n - 2) * s];
        }
        double g[CK], up[CK];
        regamma_u<NO, ORDER, CK>(cb, n, chk[(cb / CK) * gs], chku[(cb / CK) * gs], ph, xInt, df,
                                 W0, W1, S, ts, c, m, dt, inv_dt, bcl_on, bcl, g, up);
#pragma unroll
        for (int q = CK - 1; q >= 0; --q) {
            const int jr = cb + q;
            if (jr <= n - 2) {
                const double v = up[q] - g[q] * u;
                p[jr * s] = v;
                u = v;
            }
        }
    }
}

// Contiguous-axis tile of the recompute variant: staging chunks of C = CK
// columns; forward loads only, backward loads, recomputes and stores.
template <int NO, int ORDER, int CK, int C>
__device__ __forceinline__ void solve_tile_staged_ru(
    double *__restrict__ tile, int rows_active, int n, long long ld,
    const double *__restrict__ xInt, const double *__restrict__ df,
    const double *__restrict__ W0, const double *__restrict__ W1,
    const double *__restrict__ S,
    const double *c, const double *m, double dt, double inv_dt,
    bool bcf_on, double bcf, bool bcl_on, double bcl, int inj_row, double inj,
    double *__restrict__ chk, int col, double *__restrict__ buf, int lane)
{
    static_assert(C == CK, "one checkpoint chunk per staging chunk");
    constexpr int LD = C + 1;
    double *mine = buf + lane * LD;
    double *chku = chk + (long long)col * 64;
    const int nch = (n + C - 1) / C;
    double bet = 0., u = 0., pa = 0., pc = 0.;
    for (int k = 0; k < nch; ++k) {
        const int j0 = k * C;
        const int cw = n - j0 < C ? n - j0 : C;
        {
            double v[C];
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const int e = q * 64 + lane, row = e / C, cl = e - row * C;
                const bool ok = row < rows_active && cl < cw;
                v[q] = tile[ok ? (long long)row * ld + j0 + cl : 0];
                v[q] = ok ? v[q] : 0.;
            }
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const int e = q * 64 + lane, row = e / C, cl = e - row * C;
                buf[row * LD + cl] = v[q];
            }
        }
        wave_sync();
        if (k == 0 && inj_row >= 0) mine[inj_row] = mine[inj_row] + inj;
        for (int cc = 0; cc < cw; ++cc) {
            const int j = j0 + cc;
            const double pj = mine[cc];
            if (j == 0) {
                row0<NO, ORDER>(pj, xInt, df, W0, W1, S, c, m, dt, inv_dt, bcf_on, bcf, bet, u,
                                pa, pc);
                chk[0] = bet;
                chku[0] = u;
            } else {
                fwd_core<NO, ORDER>(j, n, pj, xInt, df, W0, W1, S, 1, c, m, dt, inv_dt, bcl_on,
                                    bcl, bet, u, pa, pc);
                if (j % CK == 0) {
                    chk[(j / CK) * 64] = bet;
                    chku[(j / CK) * 64] = u;
                }
            }
        }
        wave_sync();
    }
    const double ulast = u;
    for (int k = nch - 1; k >= 0; --k) {
        const int j0 = k * C;
        const int cw = n - j0 < C ? n - j0 : C;
        {
            double v[C];
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const int e = q * 64 + lane, row = e / C, cl = e - row * C;
                const bool ok = row < rows_active && cl < cw;
                v[q] = tile[ok ? (long long)row * ld + j0 + cl : 0];
                v[q] = ok ? v[q] : 0.;
            }
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const int e = q * 64 + lane, row = e / C, cl = e - row * C;
                buf[row * LD + cl] = v[q];
            }
        }
        wave_sync();
        if (k == 0 && inj_row >= 0) mine[inj_row] = mine[inj_row] + inj;
        if (j0 + cw - 1 == n - 1) mine[n - 1 - j0] = ulast;
        const int cb = j0;
        if (cb <= n - 2) {
            double ph[CK];
#pragma unroll
            for (int q = 0; q < CK; ++q) ph[q] = mine[q < cw ? q : 0];
            double g[CK], up[CK];
            regamma_u<NO, ORDER, CK>(cb, n, chk[(cb / CK) * 64], chku[(cb / CK) * 64], ph, xInt,
                                     df, W0, W1, S, 1, c, m, dt, inv_dt, bcl_on, bcl, g, up);
#pragma unroll
            for (int q = CK - 1; q >= 0; --q) {
                const int jr = cb + q;
                if (jr <= n - 2) {
                    const double v = up[q] - g[q] * u;
                    mine[q] = v;
                    u = v;
                }
            }
        }
        wave_sync();
        {
            double v[C];
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const int e = q * 64 + lane, row = e / C, cl = e - row * C;
                v[q] = buf[row * LD + cl];
            }
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const int e = q * 64 + lane, row = e / C, cl = e - row * C;
                if (row < rows_active && cl < cw) tile[(long long)row * ld + j0 + cl] = v[q];
            }
        }
        wave_sync();
    }
}

// Boundary terms and injection row of one line of axis `a`.
template <int D, int ORDER>
struct LineSetup {
    double c[D > 1 ? D - 1 : 1];
    bool bcf_on, bcl_on;
    double bcf, bcl;
    int inj_row;
};

template <int D, int ORDER>
__device__ __forceinline__ void line_setup(LineSetup<D, ORDER> &L, long long line, int n, int a,
                                           const double *__restrict__ sx, const double *mig,
                                           double nu, double gam_a, double h_a, double dx0,
                                           double dxl, bool first_axis, int injm, int chrom = 0)
{
    constexpr int NO = D - 1;
    int idx[NO > 0 ? NO : 1];
    long long rem = line;
#pragma unroll
    for (int q = NO - 1; q >= 0; --q) {
        idx[q] = (int)(rem % n);
        rem /= n;
    }
    bool all0 = true, all1 = true, first = true, last = true;
    int nones = 0, onepos = -1, nzero = 0;
#pragma unroll
    for (int q = 0; q < NO; ++q) {
        L.c[q] = sx[idx[q]];
        all0 = all0 && (L.c[q] == 0);
        all1 = all1 && (L.c[q] == 1);
        first = first && (idx[q] == 0);
        last = last && (idx[q] == n - 1);
        if (idx[q] == 1) { ++nones; onepos = q; }
        if (idx[q] == 0) ++nzero;
    }
    L.bcf_on = false;
    L.bcl_on = false;
    L.bcf = 0.;
    L.bcl = 0.;
    if constexpr (ORDER == DH_ORDER_TV) {
        // dadi/integration2D.c:67-70 and its 3-5D/1D twins
        if (all0) {
            const double Mf = d_M<NO>(L.c, mig, sx[0], d_sel(sx[0], gam_a, h_a, chrom));
            if (Mf <= 0) { L.bcf_on = true; L.bcf = (0.5 / nu - Mf) * 2. / dx0; }
        }
        if (all1) {
            const double Ml = d_M<NO>(L.c, mig, sx[n - 1], d_sel(sx[n - 1], gam_a, h_a, chrom));
            if (Ml >= 0) { L.bcl_on = true; L.bcl = -(-0.5 / nu - Ml) * 2. / dxl; }
        }
    } else {
        // dadi/Integration.py:1089-1092: corner lines only, M at x0 / x_{n-1}
        if (first) {
            double cc[NO > 0 ? NO : 1];
#pragma unroll
            for (int q = 0; q < NO; ++q) cc[q] = sx[0];
            const double M0 = d_M<NO>(cc, mig, sx[0], d_sel(sx[0], gam_a, h_a, chrom));
            if (M0 <= 0) { L.bcf_on = true; L.bcf = (0.5 / nu - M0) * 2 / dx0; }
        }
        if (last) {
            double cc[NO > 0 ? NO : 1];
#pragma unroll
            for (int q = 0; q < NO; ++q) cc[q] = sx[n - 1];
            const double M1 = d_M<NO>(cc, mig, sx[n - 1], d_sel(sx[n - 1], gam_a, h_a, chrom));
            if (M1 >= 0) { L.bcl_on = true; L.bcl = -(-0.5 / nu - M1) * 2 / dxl; }
        }
    }
    // mutation injection (dadi/Integration.py:70-145) happens before the
    // first sweep of a step; each injected point lies on its own line.
    L.inj_row = -1;
    if (first_axis) {
        if (nzero == NO && ((injm >> a) & 1)) {
            L.inj_row = 1;
        } else if (nones == 1 && nzero == NO - 1) {
            const int pop = onepos < a ? onepos : onepos + 1;
            if ((injm >> pop) & 1) L.inj_row = 0;
        }
    }
}

// Per-axis 1D coefficient vectors for one segment:
// W0[i] = V_i/(2 dx_i), W1[i] = V_{i+1}/(2 dx_i), S[i] = sel(xInt_i).
template <int D>
__device__ __forceinline__ void build_axis_tables(const dh_segment *sp, int a, int n,
                                                  const double *sx, const double *sxi,
                                                  const double *sdx, double *W0, double *W1,
                                                  double *S, int tid, int nthr,
                                                  double *VI = nullptr)
{
    const double nu = sp->nu[a], g = sp->gamma[a], h = sp->h[a];
    const double bp = sp->beta_p, br = sp->beta_r, bq = sp->beta_q;
    const int chrom = sp->chrom;
    for (int i = tid; i < n - 1; i += nthr) {
        double vi = d_vfunc(sx[i], nu), vj = d_vfunc(sx[i + 1], nu);
        if (D == 1) {  // Vfunc_beta, dadi/integration_shared.c:8-10 (br = 1), or
                       // the X-chromosome V, dadi/Integration.py:1244-1245
            vi = vi * bp * br / bq;
            vj = vj * bp * br / bq;
        }
        W0[i] = vi / (2 * sdx[i]);
        W1[i] = vj / (2 * sdx[i]);
        S[i] = d_sel(sxi[i], g, h, chrom);
        if (VI) {   // V at the interval midpoint, for the delta trick
            double v = d_vfunc(sxi[i], nu);
            if (D == 1) v = v * bp * br / bq;
            VI[i] = v;
        }
    }
}

// dx, xInt, dfactor: dadi/integration_shared.c:32-50
__device__ __forceinline__ void build_grid_tables(const double *__restrict__ xx, int n,
                                                  double *sx, double *sxi, double *sdf,
                                                  double *sdx, int tid, int nthr)
{
    for (int i = tid; i < n; i += nthr) sx[i] = xx[i];
    __syncthreads();
    for (int i = tid; i < n - 1; i += nthr) {
        sdx[i] = sx[i + 1] - sx[i];
        sxi[i] = 0.5 * (sx[i + 1] + sx[i]);
    }
    __syncthreads();
    for (int i = tid; i < n; i += nthr) {
        double v;
        if (i == 0) v = 2. / sdx[0];
        else if (i == n - 1) v = 2. / sdx[n - 2];
        else v = 2. / (sdx[i] + sdx[i - 1]);
        sdf[i] = v;
    }
}

// First element and row stride of line `line` of axis `a` of a phi whose last
// axis is stored with row pitch ld >= n (C order otherwise).  Lines enumerate
// the other axes in C order.
template <int D>
__device__ __forceinline__ void line_addr(int n, long long ld, int a, long long line,
                                          long long &off, long long &stride)
{
    if (a == D - 1) {
        off = line * ld;
        stride = 1;
        return;
    }
    long long mid = 1;   // lines per index of axis a among the axes a+1 .. D-2
    for (int q = a + 1; q < D - 1; ++q) mid *= n;
    const long long cnt = mid * n, hi = line / cnt, lo = line - hi * cnt;
    const long long lq = lo / n;
    stride = mid * ld;
    off = hi * n * stride + lq * ld + (lo - lq * n);
}

// All lines of axis `a`, tiles of 64 lines per wave.  GG: gamcol is this
// lane's gamma column in the global workspace (rows 64 doubles apart).
template <int D, int ORDER, int CK, int RU, int GG, int DJ = 0, int BP = PF, bool IB = false>
__device__ __forceinline__ void sweep_tiles(double *__restrict__ phi, int n, long long ld, int a,
                                            bool first_axis,
                                            int injm, const dh_segment *sp,
                                            const double *sx, const double *sxi,
                                            const double *sdf, const double *sdx,
                                            const double *W0, const double *W1, const double *S,
                                            double dt, double inv_dt, double *gamcol,
                                            double *stage, long long tile0, long long tstep,
                                            int lane, DeljTabs dj = {})
{
    constexpr int NO = D - 1;
    static_assert(!DJ || (CK == 0 && !RU && !GG), "the delta trick uses the plain Thomas path");
    long long nsys = 1;
    for (int q = 0; q < NO; ++q) nsys *= n;
    double mig[NO > 0 ? NO : 1];
#pragma unroll
    for (int q = 0; q < NO; ++q) mig[q] = sp->mig[a][q];
    const double nu = sp->nu[a], ga = sp->gamma[a], ha = sp->h[a];
    const double inj = sp->inject;
    if constexpr ((CK != 0 || GG) && D >= 2) {
        if (a == D - 1) {
            // contiguous axis: whole-wave tiles staged through LDS
            for (long long t = tile0; t * 64 < nsys; t += tstep) {
                const long long L0 = t * 64;
                const int rows = nsys - L0 < 64 ? (int)(nsys - L0) : 64;
                const long long line = L0 + lane;
                LineSetup<D, ORDER> L;
                line_setup<D, ORDER>(L, line < nsys ? line : nsys - 1, n, a, sx, mig, nu, ga, ha,
                                     sdx[0], sdx[n - 2], first_axis, injm, sp->chrom);
                if (line >= nsys) L.inj_row = -1;
                if constexpr (GG)
                    solve_tile_staged_gg<NO, ORDER, STAGE_C>(
                        phi + L0 * ld, rows, n, ld, sxi, sdf, W0, W1, S, L.c, mig, dt, inv_dt,
                        L.bcf_on, L.bcf, L.bcl_on, L.bcl, L.inj_row, inj, gamcol, stage, lane);
                else if constexpr (RU)
                    solve_tile_staged_ru<NO, ORDER, CK, CK>(
                        phi + L0 * ld, rows, n, ld, sxi, sdf, W0, W1, S, L.c, mig, dt, inv_dt,
                        L.bcf_on, L.bcf, L.bcl_on, L.bcl, L.inj_row, inj, gamcol,
                        (n - 1) / CK + 1, stage, lane);
                else
                    solve_tile_staged<NO, ORDER, CK, STAGE_C>(
                        TileRows<STAGE_C>{phi + L0 * ld, rows, ld}, n, sxi, sdf, W0, W1, S, L.c, mig, dt, inv_dt,
                        L.bcf_on, L.bcf, L.bcl_on, L.bcl, L.inj_row, inj, gamcol, stage, lane);
            }
            return;
        }
    }
    for (long long t = tile0; t * 64 < nsys; t += tstep) {
        const long long line = t * 64 + lane;
        if (line < nsys) {
            LineSetup<D, ORDER> L;
            line_setup<D, ORDER>(L, line, n, a, sx, mig, nu, ga, ha, sdx[0], sdx[n - 2],
                                 first_axis, injm, sp->chrom);
            long long off, s;
            line_addr<D>(n, ld, a, line, off, s);
            double *p = phi + off;
            if constexpr (CK == 0)
                solve_line<NO, ORDER, DJ, BP, IB>(p, s, n, sxi, sdf, W0, W1, S, 1, L.c, mig, dt, inv_dt,
                                          L.bcf_on, L.bcf, L.bcl_on, L.bcl, L.inj_row, inj,
                                          gamcol, 64, dj);
            else if constexpr (RU)
                solve_line_ru<NO, ORDER, CK>(p, s, n, sxi, sdf, W0, W1, S, 1, L.c, mig, dt,
                                             inv_dt, L.bcf_on, L.bcf, L.bcl_on, L.bcl, L.inj_row,
                                             inj, gamcol, 64, (n - 1) / CK + 1);
            else
                solve_line_ck<NO, ORDER, CK>(p, s, n, sxi, sdf, W0, W1, S, 1, L.c, mig, dt,
                                             inv_dt, L.bcf_on, L.bcf, L.bcl_on, L.bcl, L.inj_row,
                                             inj, gamcol, 64);
        }
    }
}

// ---------------------------------------------------------------------------
// Kernel 1: persistent, one workgroup per instance (any D), whole schedule.
// LDS: x, xInt, dfactor, dx (4n) | per axis W0,W1,S (3Dn) | gamma (waves*64*n)
// ---------------------------------------------------------------------------
template <int D, int ORDER, int CK, int RU, int GG, int DJ = 0>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(DH_WPE, 8)))
k_integrate_nd(dh_instance *__restrict__ tab, int ninst, int *__restrict__ counter,
               long long max_steps, double *__restrict__ gws, int nmax, long long max_ticks)
{
    extern __shared__ double lds[];
    __shared__ int s_next, s_stop;
    const int nthr = blockDim.x, tid = threadIdx.x;
    const int nw = nthr >> 6, wave = tid >> 6, lane = tid & 63;
    for (;;) {
        __syncthreads();
        if (tid == 0) s_next = atomicAdd(counter, 1);
        __syncthreads();
        const int id = __builtin_amdgcn_readfirstlane(s_next);
        if (id >= ninst) break;
        dh_instance *ip = tab + id;
        const int n = ip->n;
        double *sx = lds, *sxi = sx + n, *sdf = sxi + n, *sdx = sdf + n;
        double *tabs = sdx + n;
        constexpr int TPA = LDS_TABLES_PER_AXIS + DJ;   // + V at midpoints (delta trick)
        // per-lane scratch doubles: the gamma column, or (bet[, u']) checkpoints
        const int col = CK ? ((n - 1) / (CK ? CK : 1) + 1) * (RU ? 2 : 1) : n;
        const int sc = RU ? CK : STAGE_C;
        double *gamcol, *stage;
        if constexpr (GG) {
            gamcol = gws + ((long long)blockIdx.x * nw + wave) * 64 * nmax + lane;
            stage = tabs + TPA * D * n + (long long)wave * 64 * (STAGE_C + 1);
        } else {
            gamcol = tabs + TPA * D * n + (long long)wave * 64 * col + lane;
            stage = tabs + TPA * D * n + (long long)nw * 64 * col +
                    (long long)wave * 64 * (sc + 1);
        }
        build_grid_tables(gbl(ip->xx_dev), n, sx, sxi, sdf, sdx, tid, nthr);
        long long seg = ip->cur_seg, step = ip->cur_step;
        const long long nseg = ip->nseg;
        long long budget = max_steps > 0 ? max_steps : LLONG_MAX;
        const int frozen = ip->frozen, injm = ip->inject;
        double *phi = gbl(ip->phi_dev);
        const long long ld = ip->ld > 0 ? ip->ld : n;
        // time budget: thread 0 reads the clock, the whole group follows its
        // decision (read between two barriers)
        const long long t0 = max_ticks > 0 ? (long long)wall_clock64() : 0;
        bool out_of_time = false;
        while (seg < nseg && budget > 0 && !out_of_time) {
            const dh_segment *sp = gbl(ip->seg_dev) + seg;
            __syncthreads();
            for (int a = 0; a < D; ++a) {
                if ((frozen >> a) & 1) continue;
                double *W0 = tabs + TPA * a * n;
                build_axis_tables<D>(sp, a, n, sx, sxi, sdx, W0, W0 + n, W0 + 2 * n, tid, nthr,
                                     DJ ? W0 + 3 * n : nullptr);
            }
            __syncthreads();
            const double dt = sp->dt, inv_dt = 1. / dt;
            long long nst = sp->nsteps - step;
            if (nst > budget) nst = budget;
            for (long long k = 0; k < nst; ++k) {
                bool first = true;
                for (int a = 0; a < D; ++a) {
                    if ((frozen >> a) & 1) continue;
                    const double *W0 = tabs + TPA * a * n;
                    DeljTabs dj;
                    if constexpr (DJ) dj = DeljTabs{W0 + 3 * n, sdx};
                    sweep_tiles<D, ORDER, CK, RU, GG, DJ>(phi, n, ld, a, first, injm, sp, sx, sxi,
                                                          sdf, sdx, W0, W0 + n, W0 + 2 * n, dt,
                                                          inv_dt, gamcol, stage, wave, nw, lane,
                                                          dj);
                    __syncthreads();
                    first = false;
                }
                if (max_ticks > 0) {
                    if (tid == 0) s_stop = (long long)wall_clock64() - t0 >= max_ticks;
                    __syncthreads();
                    if (s_stop) {
                        nst = k + 1;
                        out_of_time = true;
                        break;
                    }
                }
            }
            budget -= nst;
            step += nst;
            if (step >= sp->nsteps) {
                ++seg;
                step = 0;
            }
        }
        if (tid == 0) {
            ip->cur_seg = seg;
            ip->cur_step = step;
            ip->status = seg >= nseg ? 1 : 0;
        }
    }
}

// ---------------------------------------------------------------------------
// Kernel 1b: persistent, several instances ("slots") per workgroup, stepped in
// lockstep so that their lines fill the waves (a 2-population phi on 100
// points has 100 lines per sweep: alone it leaves 28 of 128 lanes idle, three
// slots fill 300 of 320).  All instances of a launch share D, order, n and the
// grid.  Each slot runs its own schedule; a finished slot takes the next
// instance from the queue.  Lines of a slot are padded to a multiple of 8 so
// that every staging row group of the contiguous axis lies in one instance.
// Checkpointed Thomas scratch (bet every 8 rows) only.
// LDS: x, xInt, dfactor, dx (4n) | per slot and axis W0,W1,S (3*S*D*n) |
//      per wave: checkpoints (64*((n-1)/8+1)) and staging (64*(STAGE_C+1)).
// ---------------------------------------------------------------------------
constexpr int DH_PACK_MAX = 16;

// Pointers in the slot table are kept as global-address-space pointers so
// that phi and schedule accesses through them compile to global memory
// instructions.
typedef __attribute__((address_space(1))) double *gdptr;
typedef const __attribute__((address_space(1))) dh_segment *gsptr;

struct PackSlot {
    gdptr phi;
    gsptr sp;
    long long seg, step, nseg, budget;
    double dt, inv_dt;
    int id, frozen, injm, first_axis, live, rebuild, done;
};

// Points slot `ps` at segment `seg` of instance `ip` (skipping empty ones);
// returns false when the schedule is exhausted.
__device__ inline bool pack_enter_segment(PackSlot &ps, const dh_instance *ip)
{
    const gsptr segs = (gsptr)gbl(ip->seg_dev);
    while (ps.seg < ps.nseg && segs[ps.seg].nsteps - ps.step <= 0) {
        ++ps.seg;
        ps.step = 0;
    }
    if (ps.seg >= ps.nseg) return false;
    ps.sp = segs + ps.seg;
    ps.dt = ps.sp->dt;
    ps.inv_dt = 1. / ps.dt;
    ps.rebuild = 1;
    return true;
}

template <int D, int ORDER>
__device__ __forceinline__ void sweep_packed(int a, const PackSlot *slot, int S, int n,
                                             long long ld, long long L, long long Lp,
                                             const double *sx,
                                             const double *sxi, const double *sdf,
                                             const double *sdx, const double *atab,
                                             double *gamcol, double *stage, int wave, int nw,
                                             int lane)
{
    constexpr int NO = D - 1, CK = 8, C = STAGE_C;
    static_assert(8 % (64 / C) == 0, "row groups must not straddle the 8-line slot padding");
    const long long span = (long long)S * Lp;
    const int ntiles = (int)((span + 63) / 64);
    for (int t = wave; t < ntiles; t += nw) {
        const long long G = (long long)t * 64 + lane;
        int k = (int)(G / Lp);
        if (k > S - 1) k = S - 1;
        const long long line = G - (long long)k * Lp;
        const PackSlot &ps = slot[k];
        const bool act = G < span && line < L && ps.live && !((ps.frozen >> a) & 1);
        const dh_segment *sp = (const dh_segment *)ps.sp;
        double mig[NO > 0 ? NO : 1];
#pragma unroll
        for (int q = 0; q < NO; ++q) mig[q] = sp->mig[a][q];
        const double nu = sp->nu[a], ga = sp->gamma[a], ha = sp->h[a], inj = sp->inject;
        const double dt = ps.dt, inv_dt = ps.inv_dt;
        const double *W0 = atab + ((long long)k * D + a) * LDS_TABLES_PER_AXIS * n;
        LineSetup<D, ORDER> Ls;
        line_setup<D, ORDER>(Ls, act ? line : 0, n, a, sx, mig, nu, ga, ha, sdx[0], sdx[n - 2],
                             a == ps.first_axis, ps.injm, sp->chrom);
        if (!act) Ls.inj_row = -1;
        if (a == D - 1) {
            PackedRows<C> R;
            R.ld = ld;
            bool any = false;
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const long long Gq = (long long)t * 64 + q * (64 / C);
                int kq = (int)(Gq / Lp);
                if (kq > S - 1) kq = S - 1;
                const long long lq = Gq - (long long)kq * Lp;
                const PackSlot &pq = slot[kq];
                const bool use = Gq < span && lq < L && pq.live && !((pq.frozen >> a) & 1);
                const long long rows = L - lq < 64 / C ? L - lq : 64 / C;
                R.valid[q] = use ? (int)rows : 0;
                R.base[q] = (double *)pq.phi + (use ? lq * ld : 0);
                any = any || use;
            }
            if (!any) continue;
            solve_tile_staged<NO, ORDER, CK, C>(R, n, sxi, sdf, W0, W0 + n, W0 + 2 * n, Ls.c, mig,
                                                dt, inv_dt, Ls.bcf_on, Ls.bcf, Ls.bcl_on, Ls.bcl,
                                                Ls.inj_row, inj, gamcol, stage, lane);
        } else if (act) {
            long long off, s;
            line_addr<D>(n, ld, a, line, off, s);
            double *p = (double *)ps.phi + off;
            solve_line_ck<NO, ORDER, CK>(p, s, n, sxi, sdf, W0, W0 + n, W0 + 2 * n, 1, Ls.c, mig,
                                         dt, inv_dt, Ls.bcf_on, Ls.bcf, Ls.bcl_on, Ls.bcl,
                                         Ls.inj_row, inj, gamcol, 64);
        }
    }
}

template <int D, int ORDER>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(DH_WPE, 8)))
k_integrate_packed(dh_instance *__restrict__ tab, int ninst, int *__restrict__ counter,
                   long long max_steps, int S, int n)
{
    extern __shared__ double lds[];
    __shared__ PackSlot slot[DH_PACK_MAX];
    __shared__ int s_live;
    const int nthr = blockDim.x, tid = threadIdx.x;
    const int nw = nthr >> 6, wave = tid >> 6, lane = tid & 63;
    long long L = 1;
    for (int q = 0; q < D - 1; ++q) L *= n;
    const long long Lp = (L + 7) & ~7LL;
    const long long ld = tab[0].ld > 0 ? tab[0].ld : n;   // shared by every instance
    double *sx = lds, *sxi = sx + n, *sdf = sxi + n, *sdx = sdf + n;
    double *atab = sdx + n;
    const int col = (n - 1) / 8 + 1;
    double *scr = atab + (long long)S * D * LDS_TABLES_PER_AXIS * n;
    double *gamcol = scr + (long long)wave * 64 * col + lane;
    double *stage = scr + (long long)nw * 64 * col + (long long)wave * 64 * (STAGE_C + 1);
    build_grid_tables(gbl(tab[0].xx_dev), n, sx, sxi, sdf, sdx, tid, nthr);
    if (tid < S) {
        // an idle slot still points at valid memory: its lanes compute on
        // padding and never store
        PackSlot &ps = slot[tid];
        ps.phi = (gdptr)gbl(tab[0].phi_dev);
        ps.sp = (gsptr)gbl(tab[0].seg_dev);
        ps.dt = 1.;
        ps.inv_dt = 1.;
        ps.live = 0;
        ps.rebuild = 0;
        ps.done = 0;
        ps.id = -1;
        ps.frozen = 0;
        ps.injm = 0;
        ps.first_axis = 0;
    }
    const long long budget0 = max_steps > 0 ? max_steps : LLONG_MAX;
    for (;;) {
        __syncthreads();
        if (tid == 0) {
            int live = 0;
            for (int k = 0; k < S; ++k) {
                PackSlot &ps = slot[k];
                while (!ps.live && !ps.done) {
                    const int id = atomicAdd(counter, 1);
                    if (id >= ninst) {
                        ps.done = 1;
                        break;
                    }
                    dh_instance *ip = tab + id;
                    ps.id = id;
                    ps.seg = ip->cur_seg;
                    ps.step = ip->cur_step;
                    ps.nseg = ip->nseg;
                    ps.budget = budget0;
                    ps.frozen = ip->frozen;
                    ps.injm = ip->inject;
                    int fa = 0;
                    while (fa < D && ((ps.frozen >> fa) & 1)) ++fa;
                    ps.first_axis = fa;
                    if (fa == D) {   // every axis frozen: the schedule is the identity
                        ps.seg = ps.nseg;
                        ps.step = 0;
                    }
                    if (fa < D && pack_enter_segment(ps, ip)) {
                        ps.phi = (gdptr)gbl(ip->phi_dev);
                        ps.live = 1;
                    } else {
                        ip->cur_seg = ps.seg;
                        ip->cur_step = ps.step;
                        ip->status = ps.seg >= ps.nseg ? 1 : 0;
                    }
                }
                live += ps.live;
            }
            s_live = live;
        }
        __syncthreads();
        if (s_live == 0) break;
        for (int k = 0; k < S; ++k) {
            if (!slot[k].rebuild) continue;
            const dh_segment *sp = (const dh_segment *)slot[k].sp;
            for (int a = 0; a < D; ++a) {
                if ((slot[k].frozen >> a) & 1) continue;
                double *W0 = atab + ((long long)k * D + a) * LDS_TABLES_PER_AXIS * n;
                build_axis_tables<D>(sp, a, n, sx, sxi, sdx, W0, W0 + n, W0 + 2 * n, tid, nthr);
            }
        }
        __syncthreads();
        if (tid < S) slot[tid].rebuild = 0;
        for (int a = 0; a < D; ++a) {
            sweep_packed<D, ORDER>(a, slot, S, n, ld, L, Lp, sx, sxi, sdf, sdx, atab, gamcol, stage,
                                   wave, nw, lane);
            __syncthreads();
        }
        if (tid < S && slot[tid].live) {
            PackSlot &ps = slot[tid];
            dh_instance *ip = tab + ps.id;
            ++ps.step;
            --ps.budget;
            bool more = true;
            if (ps.step >= ps.sp->nsteps) {
                ++ps.seg;
                ps.step = 0;
                more = pack_enter_segment(ps, ip);
            }
            if (!more || ps.budget <= 0) {
                ip->cur_seg = ps.seg;
                ip->cur_step = ps.step;
                ip->status = ps.seg >= ps.nseg ? 1 : 0;
                ps.live = 0;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Kernel 1c: 2-population integrations with phi resident in registers
// (n from 33 to 104 points, in kernels for 8 consecutive n each).
// One workgroup of ceil(NP/64) waves per integration, pulled from the queue.
// Lane L holds line L of the axis being swept: the column phi[., L] for the x
// sweep, the row phi[L, .] for the y sweep (the other coordinate of both
// lines is x_L: the two axes share the grid).  phi is read from memory when
// the workgroup takes an integration up and written back when it puts it
// down; in between every step runs on chip.  The Thomas forward pass keeps
// u' over phi in registers and gamma_j in accumulation registers, the back
// substitution finishes the line in place, and between the sweeps the lines
// are transposed through an LDS tile.  The arithmetic is row0/fwd_core's, so
// the results are the other kernels' bits.  A kernel built for NP rows takes
// n in (NP - 8, NP]: rows below NP - 8 are known to exist and not to be the
// last, the last 8 are guarded.
// LDS: x, dx (2 NP) | per axis the coefficient tables of every row j,
//      interleaved as xInt_j, dfactor_j, W0_j, W1_j, S_j (5 NP each: the grid
//      tables are repeated per axis so that one address reaches a row's five
//      values) | transpose tile 64 x (NP + 1).
// ---------------------------------------------------------------------------
constexpr int RR_SPAN = 8;   // n values one kernel covers

template <int NP>
constexpr int rr_waves() { return (NP + 63) / 64; }

template <int NP>
constexpr size_t rr_lds_doubles() { return 12 * (size_t)NP + 64 * (size_t)(NP + 1); }

// Returns `off` unchanged, through an empty asm that also consumes `dep`: no
// instruction is emitted, but whatever is addressed through the result
// cannot be loaded before `dep` is computed.  Row j's coefficient tables are
// addressed through the u of row j - 2, so their loads (and the
// coefficient arithmetic on them) start a fixed distance ahead of the
// recurrence instead of all at once; hoisted, every row's coefficients would
// be live together, and within a segment they are even invariant across
// steps.  `off` is the byte offset of the axis' table block in LDS, so each
// table read is one LDS load with an immediate offset.

// A zero the compiler cannot see through, produced anew at each use: values
// formed from it inside the step loop (the line setup of a sweep) stay inside
// the loop instead of being hoisted out and kept live across steps.
__device__ __forceinline__ int rr_opaque_zero()
{
    int z;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z));
    return z;
}

__device__ __forceinline__ int rr_after(int off, double dep)
{
    asm volatile("" : "+v"(off) : "v"(dep));
    return off;
}

// A double held in two accumulation registers (the unified register file's
// second half), moved with v_accvgpr_write/read: gamma_j is written once in
// the forward pass and read once in the back substitution.
struct RrAcc {
    unsigned lo, hi;
};

__device__ __forceinline__ RrAcc rr_acc_put(double x)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    RrAcc r;
    asm volatile("v_accvgpr_write_b32 %0, %1" : "=a"(r.lo) : "v"((unsigned)b));
    asm volatile("v_accvgpr_write_b32 %0, %1" : "=a"(r.hi) : "v"((unsigned)(b >> 32)));
    return r;
}

__device__ __forceinline__ double rr_acc_get(const RrAcc &r)
{
    unsigned lo, hi;
    asm volatile("v_accvgpr_read_b32 %0, %1" : "=v"(lo) : "a"(r.lo));
    asm volatile("v_accvgpr_read_b32 %0, %1" : "=v"(hi) : "a"(r.hi));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Rows of a resident line kept in vector registers; the rest (the line's
// last rows) live in accumulation registers beside gamma.  About 78 vector
// registers stay free for the row arithmetic and the line's coefficients.
template <int NP>
constexpr int rr_vrows() { return NP <= 88 ? NP : 88; }

// This lane's line: rows 0..NP-1, row j in v[j] (j < PV) or a[j - PV].
// Indices are compile-time constants after unrolling, so get/set fold to a
// register access (or two accumulation-register moves).
template <int NP, int PVR = rr_vrows<NP>()>
struct RrLine {
    static constexpr int PV = PVR;
    double v[PV];
    RrAcc a[NP > PV ? NP - PV : 1];
    __device__ __forceinline__ double get(int j) const
    {
        return j < PV ? v[j < PV ? j : 0] : rr_acc_get(a[j >= PV ? j - PV : 0]);
    }
    __device__ __forceinline__ void set(int j, double x)
    {
        if (j < PV)
            v[j < PV ? j : 0] = x;
        else
            a[j >= PV ? j - PV : 0] = rr_acc_put(x);
    }
};

// One implicit half-step of this lane's line (rows 0..n-1); injection first
// (row 0 or 1 of this line, or none).
constexpr int RR_TS = 5;   // interleaved coefficient tables: stride per row
#ifndef RR_FENCE
#define RR_FENCE() __builtin_amdgcn_sched_barrier(0)
#endif
#ifndef RR_FENCE_EVERY
#define RR_FENCE_EVERY 4   // rows per scheduling block
#endif

// Grid tables of the interleaved layout (x and dx apart, xInt and dfactor in
// `blk`), with build_grid_tables' operations.
__device__ __forceinline__ void rr_grid_tables(const double *__restrict__ xx, int n, double *sx,
                                               double *sdx, double *blk, int tid, int nthr)
{
    for (int i = tid; i < n; i += nthr) sx[i] = xx[i];
    __syncthreads();
    for (int i = tid; i < n - 1; i += nthr) {
        sdx[i] = sx[i + 1] - sx[i];
        blk[i * RR_TS] = 0.5 * (sx[i + 1] + sx[i]);
    }
    __syncthreads();
    for (int i = tid; i < n; i += nthr) {
        double v;
        if (i == 0) v = 2. / sdx[0];
        else if (i == n - 1) v = 2. / sdx[n - 2];
        else v = 2. / (sdx[i] + sdx[i - 1]);
        blk[i * RR_TS + 1] = v;
    }
}

// W0, W1, S of axis a in the interleaved layout, with build_axis_tables'
// operations (2 populations: no beta factor).
__device__ __forceinline__ void rr_axis_tables(const dh_segment *sp, int a, int n,
                                               const double *sx, const double *sdx, double *blk,
                                               int tid, int nthr)
{
    const double nu = sp->nu[a], g = sp->gamma[a], h = sp->h[a];
    const int chrom = sp->chrom;
    for (int i = tid; i < n - 1; i += nthr) {
        const double vi = d_vfunc(sx[i], nu), vj = d_vfunc(sx[i + 1], nu);
        blk[i * RR_TS + 2] = vi / (2 * sdx[i]);
        blk[i * RR_TS + 3] = vj / (2 * sdx[i]);
        blk[i * RR_TS + 4] = d_sel(blk[i * RR_TS], g, h, chrom);
    }
}

// Next row's tables become this row's.
__device__ __forceinline__ void rr_take(double &xi, double &df, double &w0, double &w1,
                                        double &sl, double nxi, double ndf, double nw0,
                                        double nw1, double nsl)
{
    xi = nxi;
    df = ndf;
    w0 = nw0;
    w1 = nw1;
    sl = nsl;
}

// Quotients by bet_j, two per row, from one reciprocal.  x / b compiles to
//   y = rcp(b'); y = fma(y, fma(-b', y, 1), y) (twice); q = x' * y;
//   e = fma(-b', q, x'); result = div_fixup(div_fmas(e, y, q), b, x)
// where b' and x' come from div_scale.  div_scale rescales, and div_fmas and
// div_fixup adjust, only near the ends of the exponent range (a zero,
// denormal or non-finite operand, b beyond 2^1022, x below 2^-969, x / b
// denormal or above 2^767); elsewhere b' = b, x' = x and the result is
// fma(e, y, q).  rr_recip and rr_quot are that sequence with b' = b and
// x' = x, so y is shared by both quotients of a row.  With b in [2^-60,
// 2^100] and x in [2^-800, 2^500] every condition above holds, so the result
// is the bits of x / b.  rr_quot0 forms the residual as b q - x and
// subtracts it, fma(-(fma(b, q, -x)), y, q): for nonzero x the same value
// (round to nearest is symmetric, and a zero residual adds nothing to q), and
// for a zero x with b > 0 the zero of x's sign, as x / b gives (rr_quot gives
// +0 for x = -0).  The recurrences' numerators are checked by their binary
// exponent (rr_track: v_frexp_exp and one unsigned compare), which leaves
// zeros alone; what else the argument needs -- the coefficients and dt in
// range and every value finite -- is checked once per segment
// (rr_coeffs_in_range) and when an instance is taken up (rr_phi_in_range).
// A lane that fails marks its integration for the exactly dividing kernel.
__device__ __forceinline__ double rr_recip(double b)
{
    double y = __builtin_amdgcn_rcp(b);
    double t = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, t, y);
    t = __builtin_fma(-b, y, 1.0);
    return __builtin_fma(y, t, y);
}

__device__ __forceinline__ double rr_quot(double x, double b, double y)
{
    const double q = x * y;
    const double e = __builtin_fma(-b, q, x);
    return __builtin_fma(e, y, q);
}

__device__ __forceinline__ double rr_quot0(double x, double b, double y)
{
    const double q = x * y;
    const double e = __builtin_fma(b, q, -x);
    return __builtin_fma(-e, y, q);
}

__device__ __forceinline__ bool rr_in(double x, double lo, double hi)
{
    return (__builtin_fabs(x) >= lo) & (__builtin_fabs(x) <= hi);
}

constexpr double RR_DEN_LO = 0x1p-60, RR_DEN_HI = 0x1p100;
constexpr double RR_QUO_LO = 0x1p-900;
constexpr double RR_COEF_HI = 0x1p100;   // |a_j|, |gamma_j|, 1/dt
constexpr double RR_PHI_HI = 0x1p600;    // phi when an instance is taken up
// numerators x of the recurrences' quotients: 0, or 2^(e-1) <= |x| < 2^e
// with e = frexp_exp(x) in [RR_EXP_LO, RR_EXP_HI], i.e. |x| in [2^-800, 2^500)
constexpr int RR_EXP_LO = -799, RR_EXP_HI = 500;

// Positive and in [2^-60, 2^100] (false for NaN).
__device__ __forceinline__ bool rr_pos(double b)
{
    return (b >= RR_DEN_LO) & (b <= RR_DEN_HI);
}

// bad |= the lanes whose numerator x has its binary exponent outside
// [RR_EXP_LO, RR_EXP_HI] (one unsigned compare: an exponent below RR_EXP_LO
// wraps to a large value), folded at once into the wave mask in scalar
// registers (the empty asm).  frexp_exp gives 0 for zeros (accepted) and
// for infinities and NaN, which the segment and take-up checks keep out of
// the recurrences.
__device__ __forceinline__ void rr_track(unsigned long long &bad, double x)
{
    const unsigned e = (unsigned)(__builtin_amdgcn_frexp_exp(x) - RR_EXP_LO);
    bad |= __builtin_amdgcn_ballot_w64(e > (unsigned)(RR_EXP_HI - RR_EXP_LO));
    asm volatile("" : "+s"(bad));
}

// Loads row j's tables (after `dep`: see rr_after).
template <int NP>
__device__ __forceinline__ void rr_fetch(const double *lds, int &off, double dep, int j,
                                         double &xi, double &df, double &w0, double &w1,
                                         double &sl)
{
    off = rr_after(off, dep);
    const double *T = (const double *)((const char *)lds + off) + j * RR_TS;
    xi = T[0];
    df = T[1];
    w0 = T[2];
    w1 = T[3];
    sl = T[4];
}

// Row 0's coefficients (row0): the carry (pa, pc) of interface 0 and
// b_0 = bet_0, the boundary term by a select (which gives the bits of the
// conditional add).
template <int ORDER>
__device__ __forceinline__ double rr_row0(double xi, double df, double w0, double w1, double sl,
                                          double c, double m, double inv_dt, bool bcf_on,
                                          double bcf, double &pa, double &pc)
{
    const double M = m * (c - xi) + sl;
    pa = M * 0.5 + w0;
    pc = -M * (1 - 0.5) + w1;
    double b;
    if constexpr (ORDER == DH_ORDER_TV) {
        b = inv_dt + df * pa;
        b = bcf_on ? b + bcf : b;
    } else {
        b = df * pa;
        b = bcf_on ? b + bcf : b;
        b = b + inv_dt;
    }
    return b;
}

// Coefficients of a row j that is not the last (row_coeffs) from its tables
// and the carry (pa, pc) of interface j-1: a_j, b_j and c_j = -dfactor_j
// ctemp_j, the carry advanced to interface j.
template <int ORDER>
__device__ __forceinline__ void rr_coef_mid(double xi, double df, double w0, double w1,
                                            double sl, double c, double m, double inv_dt,
                                            double &pa, double &pc, double &a, double &b,
                                            double &cj)
{
    a = -df * pa;
    const double M = m * (c - xi) + sl;
    const double na = M * 0.5 + w0;
    const double nc = -M * (1 - 0.5) + w1;
    if constexpr (ORDER == DH_ORDER_TV)
        b = (inv_dt + df * pc) + df * na;
    else
        b = (df * na + df * pc) + inv_dt;
    cj = -df * nc;
    pa = na;
    pc = nc;
}

// Coefficients of the last row: a_j and b_j with its boundary term (a
// select, which gives the bits of the conditional add).
template <int ORDER>
__device__ __forceinline__ void rr_coef_last(double df, double inv_dt, bool bcl_on, double bcl,
                                             double pa, double pc, double &a, double &b)
{
    a = -df * pa;
    if constexpr (ORDER == DH_ORDER_TV) {
        b = inv_dt + df * pc;
        b = bcl_on ? b + bcl : b;
    } else {
        b = df * pc;
        b = bcl_on ? b + bcl : b;
        b = b + inv_dt;
    }
}

// Row j that is not the last, from its coefficients (a, b, cj) and gamma_j
// in g: bet_j = b_j - a_j gamma_j, u_j (fwd_core), gamma_{j+1} = c_j / bet_j
// into g.  FAST: the quotients by rr_recip / rr_quot (and pj / dt from
// y_dt, the reciprocal of dt), gamma_{j+1} only if fast_g (else by
// division: the last rows, where c_j = -0 at x = 1 when the advection term
// is zero); the numerators' exponents go into bad (bet_j and gamma_{j+1} are
// the same at every step of a segment: rr_coeffs_in_range checks them once).
template <int ORDER, bool FAST>
__device__ __forceinline__ void rr_solve_mid(double pj, double a, double b, double cj,
                                             double dt, double inv_dt, double y_dt, bool fast_g,
                                             double &g, double &u, unsigned long long &bad)
{
    const double bet = b - a * g;
    if constexpr (FAST) {
        double r;
        const double y = rr_recip(bet);
        if constexpr (ORDER == DH_ORDER_PRE) {
            r = inv_dt * pj;
        } else {
            rr_track(bad, pj);
            r = rr_quot0(pj, dt, y_dt);
        }
        const double x = r - a * u;
        rr_track(bad, x);
        u = rr_quot0(x, bet, y);
        g = fast_g ? rr_quot(cj, bet, y) : cj / bet;
    } else {
        const double r = (ORDER == DH_ORDER_PRE) ? inv_dt * pj : pj / dt;
        u = (r - a * u) / bet;
        g = cj / bet;
    }
}

// The last row: bet and u.
template <int ORDER>
__device__ __forceinline__ void rr_solve_last(double pj, double a, double b, double dt,
                                              double inv_dt, double g, double &u)
{
    const double bet = b - a * g;
    const double r = (ORDER == DH_ORDER_PRE) ? inv_dt * pj : pj / dt;
    u = (r - a * u) / bet;
}

// Whether a segment's coefficients of this lane's line (the axis' tables at
// T), formed with the operations of rr_sweep's rows, keep its quotients
// exact and its values finite: bet_j of every row positive and in [2^-60,
// 2^100], |a_j| and |gamma_{j+1}| at most 2^100, and every gamma_{j+1} that
// rr_sweep forms by rr_quot (rows 1..fast_rows-1) at least 2^-900 or from
// c_j = +0.  Within a segment these values are the same at every step.
template <int ORDER>
__device__ __forceinline__ bool rr_coeffs_in_range(int n, int fast_rows, const double *T,
                                                   double c, double m, double inv_dt,
                                                   bool bcf_on, double bcf, bool bcl_on,
                                                   double bcl)
{
    double pa, pc;
    const double bet0 = rr_row0<ORDER>(T[0], T[1], T[2], T[3], T[4], c, m, inv_dt, bcf_on,
                                       bcf, pa, pc);
    double g = (-T[1] * pc) / bet0;
    bool in = rr_pos(bet0) & (__builtin_fabs(g) <= RR_COEF_HI);
#pragma unroll 1
    for (int j = 1; j < n - 1; ++j) {
        const double *R = T + j * RR_TS;
        double a, b, cj;
        rr_coef_mid<ORDER>(R[0], R[1], R[2], R[3], R[4], c, m, inv_dt, pa, pc, a, b, cj);
        const double bet = b - a * g;
        in = in & rr_pos(bet) & (__builtin_fabs(a) <= RR_COEF_HI);
        if (j < fast_rows) {
            g = rr_quot(cj, bet, rr_recip(bet));
            in = in & (rr_in(g, RR_QUO_LO, RR_COEF_HI) |
                       (__builtin_bit_cast(unsigned long long, cj) == 0));
        } else {
            g = cj / bet;
            in = in & (__builtin_fabs(g) <= RR_COEF_HI);
        }
    }
    double a, b;
    rr_coef_last<ORDER>(T[(n - 1) * RR_TS + 1], inv_dt, bcl_on, bcl, pa, pc, a, b);
    return in & rr_pos(b - a * g) & (__builtin_fabs(a) <= RR_COEF_HI);
}

// Whether this lane's line of phi (rows 0..n-1) is finite with |phi| at most
// 2^600, which with the segment checks keeps every value of the recurrences
// finite until a numerator leaves its range.
template <int NP, class Line>
__device__ __forceinline__ bool rr_phi_in_range(const Line &P, int n)
{
    bool in = true;
#pragma unroll
    for (int i = 0; i < NP; ++i)
        if (i < n) in = in & (__builtin_fabs(P.get(i)) <= RR_PHI_HI);
    return in;
}

// Rows 1..rr_fast_g_rows-1 of an NP-row kernel compute gamma_{j+1} by
// rr_quot; the rest divide (row n-2 is at least NP - 9).
template <int NP>
constexpr int rr_fast_g_rows() { return NP - RR_SPAN - 1; }

// One implicit half-step of this lane's line (rows 0..n-1), the axis'
// interleaved tables at lds + off (bytes); injection first (row 0 or 1 of
// this line, or none).  The operations are those of row0 / row_coeffs /
// fwd_core (compute_abc_nobc, dadi/integration_shared.c:72-93, and
// dadi/tridiag.c:22-33), regrouped into one block per row j, pipelined:
// block j takes row j+1's tables (loaded in block j-1) and starts the loads
// of row j+2's, solves row j (bet_j = b_j - a_j gamma_j, then the two
// quotients by bet_j, u_j = (r_j - a_j u_{j-1}) / bet_j and
// gamma_{j+1} = c_j / bet_j), and forms row j+1's coefficients, which need
// neither u nor gamma, beside that chain.  A scheduling fence between blocks
// keeps each block's loads a block ahead of their use.
template <int NP, int ORDER, bool FAST>
__device__ __forceinline__ void rr_sweep(RrLine<NP> &P, RrAcc (&G)[NP], int n,
                                         const double *lds, int off, double c, double m,
                                         double dt, double inv_dt, double y_dt, bool bcf_on,
                                         double bcf, bool bcl_on, double bcl, int inj_row,
                                         double inj, unsigned long long &bad)
{
    constexpr int SURE = NP - RR_SPAN;   // rows 0..SURE exist; rows < SURE are not the last
    if (inj_row == 0) P.set(0, P.get(0) + inj);
    if (inj_row == 1) P.set(1, P.get(1) + inj);
    const double *T = (const double *)((const char *)lds + off);
    double u, pa, pc, g;
    double ca, cb, cc;   // coefficients a, b, c of the next row to solve
    // tables of the row after it (in flight)
    double nxi, ndf, nw0, nw1, nsl;
    {
        // row 0 (row0), then row 1's coefficients; row 2's tables in flight
        const double df0 = T[1];
        const double bet = rr_row0<ORDER>(T[0], df0, T[2], T[3], T[4], c, m, inv_dt, bcf_on,
                                          bcf, pa, pc);
        const double p0 = P.get(0);
        const double r = (ORDER == DH_ORDER_PRE) ? inv_dt * p0 : p0 / dt;
        u = r / bet;
        P.set(0, u);
        g = (-df0 * pc) / bet;   // gamma_1 = c_0 / bet_0
        G[1] = rr_acc_put(g);
        rr_coef_mid<ORDER>(T[RR_TS], T[RR_TS + 1], T[RR_TS + 2], T[RR_TS + 3], T[RR_TS + 4], c,
                           m, inv_dt, pa, pc, ca, cb, cc);
        rr_fetch<NP>(lds, off, u, 2, nxi, ndf, nw0, nw1, nsl);
    }
    // rows 1 .. SURE-1: present and not the last; row j+1 is not the last
    // either, except possibly for j = SURE-1
#pragma unroll
    for (int j = 1; j < SURE; ++j) {
        if (j % RR_FENCE_EVERY == 0) RR_FENCE();
        double xi, df, w0, w1, sl;
        rr_take(xi, df, w0, w1, sl, nxi, ndf, nw0, nw1, nsl);
        if (j + 2 < NP) rr_fetch<NP>(lds, off, u, j + 2, nxi, ndf, nw0, nw1, nsl);
        rr_solve_mid<ORDER, FAST>(P.get(j), ca, cb, cc, dt, inv_dt, y_dt,
                                  j < rr_fast_g_rows<NP>(), g, u, bad);
        P.set(j, u);
        G[j + 1] = rr_acc_put(g);
        if (j + 1 < SURE || j + 1 < n - 1)
            rr_coef_mid<ORDER>(xi, df, w0, w1, sl, c, m, inv_dt, pa, pc, ca, cb, cc);
        else
            rr_coef_last<ORDER>(df, inv_dt, bcl_on, bcl, pa, pc, ca, cb);
    }
    // rows SURE .. NP-1: row SURE is present, the others if j < n; the last
    // present row takes the boundary term
#pragma unroll
    for (int j = SURE; j < NP; ++j) {
        if (j % RR_FENCE_EVERY == 0) RR_FENCE();
        double xi, df, w0, w1, sl;
        rr_take(xi, df, w0, w1, sl, nxi, ndf, nw0, nw1, nsl);
        if (j + 2 < NP) rr_fetch<NP>(lds, off, u, j + 2, nxi, ndf, nw0, nw1, nsl);
        if (j == SURE || j < n) {
            if (j < n - 1) {
                rr_solve_mid<ORDER, FAST>(P.get(j), ca, cb, cc, dt, inv_dt, y_dt, false, g, u,
                                          bad);
                P.set(j, u);
                if (j + 1 < NP) G[j + 1] = rr_acc_put(g);
                if (j + 1 < n - 1)
                    rr_coef_mid<ORDER>(xi, df, w0, w1, sl, c, m, inv_dt, pa, pc, ca, cb, cc);
                else
                    rr_coef_last<ORDER>(df, inv_dt, bcl_on, bcl, pa, pc, ca, cb);
            } else {
                rr_solve_last<ORDER>(P.get(j), ca, cb, dt, inv_dt, g, u);
                P.set(j, u);
                if (j + 1 < NP) G[j + 1] = rr_acc_put(0.);   // defined on every path
            }
        } else if (j + 1 < NP) {
            G[j + 1] = rr_acc_put(0.);
        }
    }
#pragma unroll
    for (int j = NP - 2; j >= 0; --j) {
        if (j < SURE || j <= n - 2) {
            const double v = P.get(j) - rr_acc_get(G[j + 1]) * u;
            P.set(j, v);
            u = v;
        }
    }
}

// Transposes the n x n matrix whose line L lane L holds (row k of the line =
// M[k][L] -> M[L][k]) through the LDS tile (64 rows of pitch NP + 1).  With
// two waves it runs in two halves, each lane exporting rows 0..63 and then
// 64..n-1; wave 0 imports its lines after the first half, so it first keeps
// its rows 64..n-1 in X (accumulation registers, free between sweeps).
// `wave` is uniform (from readfirstlane).
template <int NP>
__device__ __forceinline__ void rr_transpose(RrLine<NP> &P, RrAcc (&X)[NP > 64 ? NP - 64 : 1],
                                             int n, int L, int wave, double *__restrict__ tile)
{
    constexpr int LDW = NP + 1;
    if constexpr (NP <= 64) {
        if (L < n) {
#pragma unroll
            for (int k = 0; k < NP; ++k)
                if (k < n) tile[k * LDW + L] = P.get(k);
        }
        wave_sync();
        if (L < n) {
#pragma unroll
            for (int k = 0; k < NP; ++k)
                if (k < n) P.set(k, tile[L * LDW + k]);
        }
        wave_sync();
    } else {
        static_assert(NP <= 128, "two waves hold at most 128 lines");
        if (L < n) {
#pragma unroll
            for (int k = 0; k < 64; ++k) tile[k * LDW + L] = P.get(k);
        }
        __syncthreads();
        // both waves fill X (only wave 0 needs it): defined on every path, X
        // is dead outside the transpose
#pragma unroll
        for (int q = 0; q < NP - 64; ++q) X[q] = rr_acc_put(P.get(64 + q));
        if (wave == 0) {
#pragma unroll
            for (int k = 0; k < NP; ++k)
                if (k < n) P.set(k, tile[L * LDW + k]);
        }
        __syncthreads();
        if (L < n) {
            if (wave == 0) {
#pragma unroll
                for (int q = 0; q < NP - 64; ++q)
                    if (64 + q < n) tile[q * LDW + L] = rr_acc_get(X[q]);
            } else {
#pragma unroll
                for (int q = 0; q < NP - 64; ++q)
                    if (64 + q < n) tile[q * LDW + L] = P.get(64 + q);
            }
        }
        __syncthreads();
        if (wave == 1 && L < n) {
#pragma unroll
            for (int k = 0; k < NP; ++k)
                if (k < n) P.set(k, tile[(L - 64) * LDW + k]);
        }
        __syncthreads();
    }
}

// Whether a lane of this wave that holds a line left the exact range.
__device__ __forceinline__ int rr_wave_out(unsigned long long okm, unsigned long long bad,
                                           int L, int n)
{
    return ((~okm | bad) & __builtin_amdgcn_ballot_w64(L < n)) != 0;
}

// FAST: the quotients of rows 1..n-2 by rr_recip / rr_quot.  A workgroup
// that finds a lane outside their exact range (every 32 steps and before it
// stores phi) leaves the instance as it took it up and sets bit 0 of its
// `reserved` word; the exactly dividing kernel (FAST false) runs only the
// instances with that bit set, from the same state, so every instance ends
// with the bits of exact division.
template <int NP, int ORDER, bool FAST>
__global__ void __launch_bounds__(64 * rr_waves<NP>()) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_integrate_rr(dh_instance *__restrict__ tab, int ninst, int *__restrict__ counter,
               long long max_steps, long long max_ticks)
{
    extern __shared__ double lds[];
    __shared__ int s_next, s_stop;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int L = tid;
    double *sx = lds, *sdx = sx + NP;
    double *blk0 = sdx + NP, *blk1 = blk0 + 5 * NP;   // per axis: xInt, dfactor, W0, W1, S
    double *tile = blk1 + 5 * NP;
    for (;;) {
        __syncthreads();
        if (tid == 0) s_next = atomicAdd(counter, 1);
        __syncthreads();
        const int id = __builtin_amdgcn_readfirstlane(s_next);
        if (id >= ninst) break;
        dh_instance *ip = tab + id;
        const int n = __builtin_amdgcn_readfirstlane(ip->n);
        if (ip->D != 2 || n <= NP - RR_SPAN || n > NP) {
            // not this kernel's shape: left untouched, reported as not done
            if (tid == 0) ip->status = -1;
            continue;
        }
        __builtin_assume(n > NP - RR_SPAN && n <= NP);
        // the fast kernel leaves instances marked for exact division to the
        // other kernel, which runs only those
        if (FAST == ((__builtin_amdgcn_readfirstlane(ip->reserved) & 1) != 0)) continue;
        rr_grid_tables(gbl(ip->xx_dev), n, sx, sdx, blk0, tid, nthr);
        __syncthreads();
        for (int i = tid; i < n; i += nthr) {
            blk1[i * RR_TS] = blk0[i * RR_TS];
            blk1[i * RR_TS + 1] = blk0[i * RR_TS + 1];
        }
        long long seg = ip->cur_seg, step = ip->cur_step;
        const long long nseg = ip->nseg;
        long long budget = max_steps > 0 ? max_steps : LLONG_MAX;
        const int frozen = ip->frozen, injm = ip->inject;
        int first_axis = 0;
        while (first_axis < 2 && ((frozen >> first_axis) & 1)) ++first_axis;
        double *phi = gbl(ip->phi_dev);
        const long long ld = ip->ld > 0 ? ip->ld : n;
        RrLine<NP> P;
        RrAcc G[NP];
        RrAcc X[NP > 64 ? NP - 64 : 1];
        if (L < n) {
#pragma unroll
            for (int i = 0; i < NP; ++i)
                if (i < n) P.set(i, phi[i * ld + L]);
        }
        // all loads issued here: their addresses need not stay live
        asm volatile("" ::: "memory");
        const long long t0 = max_ticks > 0 ? (long long)wall_clock64() : 0;
        bool out_of_time = false, failed = false;
        // FAST: the lanes whose phi and segments so far were in range (lanes
        // past the last line hold no line and count as in range), and those
        // with a numerator out of range so far (rr_track)
        unsigned long long okm = __builtin_amdgcn_ballot_w64(
            !FAST || L >= n || rr_phi_in_range<NP>(P, n));
        unsigned long long bad = 0;
        if (first_axis == 2) seg = nseg;   // every axis frozen: the identity
        while (seg < nseg && budget > 0 && !out_of_time && !failed) {
            const dh_segment *sp = gbl(ip->seg_dev) + seg;
            __syncthreads();
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                if ((frozen >> a) & 1) continue;
                rr_axis_tables(sp, a, n, sx, sdx, a ? blk1 : blk0, tid, nthr);
            }
            __syncthreads();
            const double dt = sp->dt, inv_dt = 1. / dt, inj = sp->inject;
            const double y_dt = rr_recip(dt);
            const int line = L < n ? L : 0;
            if constexpr (FAST) {
                bool in = rr_pos(dt) & (__builtin_fabs(inj) <= RR_PHI_HI);
#pragma unroll 1
                for (int a = 0; a < 2; ++a) {
                    if ((frozen >> a) & 1) continue;
                    LineSetup<2, ORDER> Ls;
                    const double mg[1] = {sp->mig[a][0]};
                    line_setup<2, ORDER>(Ls, line, n, a, sx, mg, sp->nu[a], sp->gamma[a],
                                         sp->h[a], sdx[0], sdx[n - 2], first_axis == a, injm,
                                         sp->chrom);
                    in = in & rr_coeffs_in_range<ORDER>(n, rr_fast_g_rows<NP>(),
                                                        a ? blk1 : blk0, Ls.c[0], mg[0],
                                                        inv_dt, Ls.bcf_on, Ls.bcf, Ls.bcl_on,
                                                        Ls.bcl);
                }
                okm &= __builtin_amdgcn_ballot_w64(in || L >= n);
            }
            long long nst = sp->nsteps - step;
            if (nst > budget) nst = budget;
            for (long long k = 0; k < nst; ++k) {
#pragma unroll 1
                for (int a = 0; a < 2; ++a) {
                    if (!((frozen >> a) & 1)) {
                        // the line's boundary terms, injection row and other
                        // coordinate, formed again for each sweep (through an
                        // opaque zero, so they are not kept live across steps)
                        LineSetup<2, ORDER> Ls;
                        const double mg[1] = {sp->mig[a][0]};
                        line_setup<2, ORDER>(Ls, line + rr_opaque_zero(), n, a, sx, mg,
                                             sp->nu[a], sp->gamma[a], sp->h[a], sdx[0],
                                             sdx[n - 2], first_axis == a, injm, sp->chrom);
                        const int off = (int)((const char *)(a ? blk1 : blk0) - (const char *)lds);
                        rr_sweep<NP, ORDER, FAST>(P, G, n, lds, off, Ls.c[0], mg[0], dt, inv_dt,
                                                  y_dt, Ls.bcf_on, Ls.bcf, Ls.bcl_on, Ls.bcl,
                                                  Ls.inj_row, inj, bad);
                    }
                    rr_transpose<NP>(P, X, n, L, wave, tile);
                }
                if (FAST && (k & 31) == 31 && __syncthreads_or(rr_wave_out(okm, bad, L, n))) {
                    nst = k + 1;
                    failed = true;
                    break;
                }
                if (max_ticks > 0) {
                    if (tid == 0) s_stop = (long long)wall_clock64() - t0 >= max_ticks;
                    __syncthreads();
                    if (s_stop) {
                        nst = k + 1;
                        out_of_time = true;
                        break;
                    }
                }
            }
            budget -= nst;
            step += nst;
            if (step >= sp->nsteps) {
                ++seg;
                step = 0;
            }
        }
        if (FAST && (failed || __syncthreads_or(rr_wave_out(okm, bad, L, n)))) {
            // phi and the progress are as taken up: the exact kernel redoes them
            if (tid == 0) ip->reserved |= 1;
            continue;
        }
        if (L < n) {
#pragma unroll
            for (int i = 0; i < NP; ++i)
                if (i < n) phi[i * ld + L] = P.get(i);
        }
        if (tid == 0) {
            ip->cur_seg = seg;
            ip->cur_step = step;
            ip->status = seg >= nseg ? 1 : 0;
        }
    }
}

// ---------------------------------------------------------------------------
// Kernel 1d: latency form of the register-resident 2D kernel.
// phi is held in registers as in k_integrate_rr, but within a segment every
// step solves the same tridiagonal systems, so at the segment's start each
// lane forms its line's Thomas factorisation once -- a_j, bet_j,
// y_j = rr_recip(bet_j) and gamma_{j+1}, with the operations k_integrate_rr
// uses for them at every step -- and keeps it in a global workspace laid out
// [axis][row][line] (consecutive lanes, consecutive addresses).  A step then
// runs only the right-hand-side recurrences: r_j, u_j = (r_j - a_j u_{j-1}) /
// bet_j (the shared-reciprocal quotient, as k_integrate_rr) forward and
// u_j -= gamma_{j+1} u_{j+1} back, about a quarter of k_integrate_rr's
// instructions per row, at the price of reading the records (40 bytes per
// point and sweep) from the L2.  The transpose between sweeps goes through
// one n x (n + 1) LDS tile: all lanes write, then all lanes read.  One
// workgroup per integration, one wave per SIMD: for single integrations and
// batches too small to fill the GPU.  The same exactness checks as
// k_integrate_rr; an instance that leaves their range is handed to the
// exactly dividing form (FAST false: IEEE divisions in the recurrences and
// in every gamma, the bits of k_integrate_rr's exact twin), which runs only
// the instances so marked.  NP up to 128 (two waves; the tile's LDS allows
// one workgroup per compute unit from NP = 96 on).
// ---------------------------------------------------------------------------
constexpr int RF_PF = 8;       // rows of records in flight in the forward pass
constexpr int RF_PB = 8;       // rows of gamma in flight in the back substitution

// Rows of the line in vector registers (the rest in accumulation registers):
// the record prefetch needs about 60 vector registers.
template <int NP>
constexpr int rf_vrows() { return NP <= 72 ? NP : 72; }

template <int NP>
constexpr size_t rf_lds_doubles()
{
    return (size_t)NP * (NP + 1) > 12 * (size_t)NP ? (size_t)NP * (NP + 1) : 12 * (size_t)NP;
}

// Workspace doubles of one workgroup: per axis and row, (a_j, bet_j) and
// (y_j, gamma_{j+1}) for every line, then a checkpoint of phi (NP x NP, row
// j of line L at j * NP + L).
template <int NP>
constexpr long long rf_records_doubles() { return 2LL * 2 * NP * NP * 2; }
template <int NP>
constexpr long long rf_slice_doubles() { return rf_records_doubles<NP>() + (long long)NP * NP; }

// The factorisation of line `line` for one axis, stored (if `store`) as the
// records of rows 0..n-1 at AB[j * NP + line] / YG[j * NP + line]; returns
// the checks of rr_coeffs_in_range: whether every bet_j of rows
// 1..n-2 and every gamma_{j+1} formed by rr_quot is in the range where the
// shared-reciprocal quotients give the bits of division (rr_coeffs_in_range).
// Row 0 keeps bet_0 and gamma_1 (its quotients are IEEE divisions), the last
// row a and bet.
template <int NP, int ORDER, bool FAST>
__device__ __forceinline__ bool rf_factor(double2 *__restrict__ AB, double2 *__restrict__ YG,
                                          int n, int line, bool store, const double *T,
                                          double c, double m, double inv_dt, bool bcf_on,
                                          double bcf, bool bcl_on, double bcl)
{
    double pa, pc;
    const double bet0 = rr_row0<ORDER>(T[0], T[1], T[2], T[3], T[4], c, m, inv_dt, bcf_on, bcf,
                                       pa, pc);
    double g = (-T[1] * pc) / bet0;
    if (store) {
        AB[line] = make_double2(0., bet0);
        YG[line] = make_double2(0., g);
    }
    bool in = rr_pos(bet0) & (__builtin_fabs(g) <= RR_COEF_HI);
    double ca, cb, cc;
    if (1 < n - 1)
        rr_coef_mid<ORDER>(T[RR_TS], T[RR_TS + 1], T[RR_TS + 2], T[RR_TS + 3], T[RR_TS + 4], c,
                           m, inv_dt, pa, pc, ca, cb, cc);
    else
        rr_coef_last<ORDER>(T[RR_TS + 1], inv_dt, bcl_on, bcl, pa, pc, ca, cb);
#pragma unroll 1
    for (int j = 1; j < n - 1; ++j) {
        const double bet = cb - ca * g;
        const double y = FAST ? rr_recip(bet) : 0.;
        in = in & rr_pos(bet) & (__builtin_fabs(ca) <= RR_COEF_HI);
        double gn;
        if (FAST && j < rr_fast_g_rows<NP>()) {
            gn = rr_quot(cc, bet, y);
            in = in & (rr_in(gn, RR_QUO_LO, RR_COEF_HI) |
                       (__builtin_bit_cast(unsigned long long, cc) == 0));
        } else {
            gn = cc / bet;
            in = in & (__builtin_fabs(gn) <= RR_COEF_HI);
        }
        if (store) {
            AB[(long long)j * NP + line] = make_double2(ca, bet);
            YG[(long long)j * NP + line] = make_double2(y, gn);
        }
        g = gn;
        const double *R = T + (j + 1) * RR_TS;
        if (j + 1 < n - 1)
            rr_coef_mid<ORDER>(R[0], R[1], R[2], R[3], R[4], c, m, inv_dt, pa, pc, ca, cb, cc);
        else
            rr_coef_last<ORDER>(R[1], inv_dt, bcl_on, bcl, pa, pc, ca, cb);
    }
    const double betl = cb - ca * g;
    if (store) AB[(long long)(n - 1) * NP + line] = make_double2(ca, betl);
    return in & rr_pos(betl) & (__builtin_fabs(ca) <= RR_COEF_HI);
}

// A pointer that is the same in every lane, moved into scalar registers
// (readfirstlane of both halves), as a global-address-space pointer.
__device__ __forceinline__ const __attribute__((address_space(1))) char *rf_sbase(const char *p)
{
    const unsigned long long v = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return reinterpret_cast<const __attribute__((address_space(1))) char *>(
        ((unsigned long long)hi << 32) | lo);
}

// A record of a row: the row's base (uniform) plus this lane's 32-bit byte
// offset, so the load takes its base from scalar registers and its offset
// from one vector register, with no vector address arithmetic (64-bit
// addresses formed per row were a third of the forward pass's VALU
// instructions).
__device__ __forceinline__ double2 rf_rec2(const char *row, int voff)
{
    typedef double v2d __attribute__((ext_vector_type(2)));
    const v2d v = *reinterpret_cast<const __attribute__((address_space(1))) v2d *>(
        rf_sbase(row) + (unsigned)voff);
    return make_double2(v.x, v.y);
}

__device__ __forceinline__ double rf_rec1(const char *row, int voff)
{
    return *reinterpret_cast<const __attribute__((address_space(1))) double *>(
        rf_sbase(row) + (unsigned)voff);
}

// One implicit half-step of this lane's line from its factorisation (the
// records of line L): the operations of rr_sweep's rows (row 0 and the last
// row by IEEE division, rows 1..n-2 by the shared reciprocal y_j), records
// read RF_PF rows ahead.
template <int NP, int ORDER, bool FAST>
__device__ __forceinline__ void rf_sweep(RrLine<NP, rf_vrows<NP>()> &P,
                                         const double2 *__restrict__ AB,
                                         const double2 *__restrict__ YG, int n, int L, double dt,
                                         double inv_dt, double y_dt, int inj_row, double inj,
                                         unsigned long long &bad)
{
    constexpr int SURE = NP - RR_SPAN;   // rows 0..SURE exist; rows < SURE are not the last
    constexpr size_t ROW = (size_t)NP * sizeof(double2);   // bytes per row of records
    if (inj_row == 0) P.set(0, P.get(0) + inj);
    if (inj_row == 1) P.set(1, P.get(1) + inj);
    // uniform row bases (scalar registers) plus this lane's 32-bit byte
    // offset; the offset is passed through rr_after before each load, which
    // ties the load to the recurrence (it is not hoisted) at no instruction
    // cost
    const char *ab0 = reinterpret_cast<const char *>(AB);
    const char *yg0 = reinterpret_cast<const char *>(YG);
    int voff = L * (int)sizeof(double2);
    double2 fab[RF_PF];
    double fy[RF_PF];
#pragma unroll
    for (int q = 0; q < RF_PF; ++q) {
        fab[q] = rf_rec2(ab0 + q * ROW, voff);
        fy[q] = rf_rec1(yg0 + q * ROW, voff);
    }
    double u;
    {
        const double p0 = P.get(0);
        const double r = (ORDER == DH_ORDER_PRE) ? inv_dt * p0 : p0 / dt;
        u = r / fab[0].y;
        P.set(0, u);
    }
#pragma unroll
    for (int j = 1; j < NP; ++j) {
        const double2 ab = fab[j % RF_PF];
        const double y = fy[j % RF_PF];
        if (j - 1 + RF_PF < NP) {
            // row j - 1's slot takes row j - 1 + RF_PF, issued once u_{j-1} is known
            voff = rr_after(voff, u);
            const size_t row = (size_t)(j - 1 + RF_PF) * ROW;
            fab[(j - 1) % RF_PF] = rf_rec2(ab0 + row, voff);
            fy[(j - 1) % RF_PF] = rf_rec1(yg0 + row, voff);
        }
        if (j < SURE || j < n) {
            const double pj = P.get(j);
            if (FAST && (j < SURE || j < n - 1)) {
                double r;
                if constexpr (ORDER == DH_ORDER_PRE) {
                    r = inv_dt * pj;
                } else {
                    rr_track(bad, pj);
                    r = rr_quot0(pj, dt, y_dt);
                }
                const double x = r - ab.x * u;
                rr_track(bad, x);
                u = rr_quot0(x, ab.y, y);
            } else {
                const double r = (ORDER == DH_ORDER_PRE) ? inv_dt * pj : pj / dt;
                u = (r - ab.x * u) / ab.y;
            }
            P.set(j, u);
        }
    }
    // back substitution: gamma_{j+1} is the second half of row j's (y, gamma)
    // record
    const char *g0 = yg0 + sizeof(double);
    double gb[RF_PB];
#pragma unroll
    for (int q = 0; q < RF_PB; ++q) {
        const int j = NP - 2 - q;
        gb[q] = j >= 0 ? rf_rec1(g0 + (size_t)j * ROW, voff) : 0.;
    }
#pragma unroll
    for (int j = NP - 2; j >= 0; --j) {
        const int s = (NP - 2 - j) % RF_PB;
        const double g = gb[s];
        if (j - RF_PB >= 0) {
            voff = rr_after(voff, u);
            gb[s] = rf_rec1(g0 + (size_t)(j - RF_PB) * ROW, voff);
        }
        if (j < SURE || j <= n - 2) {
            const double v = P.get(j) - g * u;
            P.set(j, v);
            u = v;
        }
    }
}

// Transposes the n x n block whose line L lane L holds (row k of the line =
// M[k][L] -> M[L][k]) through the LDS tile of pitch NP + 1: every lane
// writes its line, then every lane reads its new one.
template <int NP>
__device__ __forceinline__ void rf_transpose(RrLine<NP, rf_vrows<NP>()> &P, int n, int L,
                                             double *__restrict__ tile)
{
    constexpr int LDW = NP + 1;
    if (L < n) {
#pragma unroll
        for (int k = 0; k < NP; ++k)
            if (k < n) tile[k * LDW + L] = P.get(k);
    }
    __syncthreads();
    if (L < n) {
#pragma unroll
        for (int k = 0; k < NP; ++k)
            if (k < n) P.set(k, tile[L * LDW + k]);
    }
    __syncthreads();
}

// Where rf_run starts: phi rows at src[i * src_ld + L] (the instance's phi,
// or the checkpoint), progress (seg, step), at most `budget` steps.
struct RfStart {
    const double *src;
    long long src_ld, seg, step, budget;
};

// One instance from `at` to the end of its schedule or of the budget; phi
// and the progress are stored at the end.  FAST: every 32 steps whose checks
// pass, phi goes to the checkpoint (ckp) and `at` records it; when a check
// fails, false is returned with nothing stored (the steps since the last
// checkpoint are discarded).  The exact form always completes.
template <int NP, int ORDER, bool FAST>
__device__ __forceinline__ bool rf_run(dh_instance *ip, int n, RfStart &at, long long max_ticks,
                                       double2 *ws, double *ckp, double *lds, int *s_stop)
{
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int L = tid;
    // the tile; at a segment's start the coefficient tables (the same
    // interleaved layout as k_integrate_rr) in its place
    double *tile = lds;
    double *sx = lds, *sdx = sx + NP;
    double *blk0 = sdx + NP, *blk1 = blk0 + 5 * NP;
    long long seg = at.seg, step = at.step;
    const long long nseg = ip->nseg;
    long long budget = at.budget;
    const int frozen = ip->frozen, injm = ip->inject;
    int first_axis = 0;
    while (first_axis < 2 && ((frozen >> first_axis) & 1)) ++first_axis;
    double *phi = gbl(ip->phi_dev);
    const long long ld = ip->ld > 0 ? ip->ld : n;
    RrLine<NP, rf_vrows<NP>()> P;
    if (L < n) {
        const double *src = at.src;
#pragma unroll
        for (int i = 0; i < NP; ++i)
            if (i < n) P.set(i, src[i * at.src_ld + L]);
    }
    asm volatile("" ::: "memory");
    const long long t0 = max_ticks > 0 ? (long long)wall_clock64() : 0;
    bool out_of_time = false, failed = false;
    unsigned long long okm =
        __builtin_amdgcn_ballot_w64(!FAST || L >= n || rr_phi_in_range<NP>(P, n));
    unsigned long long bad = 0;
    if (first_axis == 2) seg = nseg;
    while (seg < nseg && budget > 0 && !out_of_time && !failed) {
        const dh_segment *sp = gbl(ip->seg_dev) + seg;
        // the tile is free between steps: the tables go there
        __syncthreads();
        rr_grid_tables(gbl(ip->xx_dev), n, sx, sdx, blk0, tid, nthr);
        __syncthreads();
        for (int i = tid; i < n; i += nthr) {
            blk1[i * RR_TS] = blk0[i * RR_TS];
            blk1[i * RR_TS + 1] = blk0[i * RR_TS + 1];
        }
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            if ((frozen >> a) & 1) continue;
            rr_axis_tables(sp, a, n, sx, sdx, a ? blk1 : blk0, tid, nthr);
        }
        __syncthreads();
        const double dt = sp->dt, inv_dt = 1. / dt, inj = sp->inject;
        const double y_dt = rr_recip(dt);
        const int line = L < n ? L : 0;
        bool in = rr_pos(dt) & (__builtin_fabs(inj) <= RR_PHI_HI);
        int inj_row[2] = {-1, -1};
#pragma unroll 1
        for (int a = 0; a < 2; ++a) {
            if ((frozen >> a) & 1) continue;
            LineSetup<2, ORDER> Ls;
            const double mg[1] = {sp->mig[a][0]};
            line_setup<2, ORDER>(Ls, line, n, a, sx, mg, sp->nu[a], sp->gamma[a], sp->h[a],
                                 sdx[0], sdx[n - 2], first_axis == a, injm, sp->chrom);
            inj_row[a] = Ls.inj_row;
            double2 *AB = ws + (long long)a * 2 * NP * NP, *YG = AB + (long long)NP * NP;
            // lanes past the last line form line 0's values and store nothing
            in = in & rf_factor<NP, ORDER, FAST>(AB, YG, n, line, L < n, a ? blk1 : blk0,
                                                 Ls.c[0], mg[0], inv_dt, Ls.bcf_on, Ls.bcf,
                                                 Ls.bcl_on, Ls.bcl);
        }
        okm &= __builtin_amdgcn_ballot_w64(!FAST || in || L >= n);
        // the tables are read: the transposes may use the tile again
        __syncthreads();
        long long nst = sp->nsteps - step;
        if (nst > budget) nst = budget;
        for (long long k = 0; k < nst; ++k) {
#pragma unroll 1
            for (int a = 0; a < 2; ++a) {
                if (!((frozen >> a) & 1)) {
                    const double2 *AB = ws + (long long)a * 2 * NP * NP;
                    rf_sweep<NP, ORDER, FAST>(P, AB, AB + (long long)NP * NP, n, line, dt,
                                              inv_dt, y_dt, inj_row[a], inj, bad);
                }
                rf_transpose<NP>(P, n, L, tile);
            }
            if (FAST && (k & 31) == 31) {
                if (__syncthreads_or(rr_wave_out(okm, bad, L, n))) {
                    nst = k + 1;
                    failed = true;
                    break;
                }
                // steps checked: the checkpoint moves here
                if (L < n) {
#pragma unroll
                    for (int i = 0; i < NP; ++i)
                        if (i < n) ckp[i * NP + L] = P.get(i);
                }
                at = RfStart{ckp, NP, seg, step + k + 1, budget - (k + 1)};
            }
            if (max_ticks > 0) {
                if (tid == 0) *s_stop = (long long)wall_clock64() - t0 >= max_ticks;
                __syncthreads();
                if (*s_stop) {
                    nst = k + 1;
                    out_of_time = true;
                    break;
                }
            }
        }
        budget -= nst;
        step += nst;
        if (step >= sp->nsteps) {
            ++seg;
            step = 0;
        }
    }
    if (FAST && (failed || __syncthreads_or(rr_wave_out(okm, bad, L, n)))) return false;
    if (L < n) {
#pragma unroll
        for (int i = 0; i < NP; ++i)
            if (i < n) phi[i * ld + L] = P.get(i);
    }
    if (tid == 0) {
        ip->cur_seg = seg;
        ip->cur_step = step;
        ip->status = seg >= nseg ? 1 : 0;
    }
    return true;
}

// FAST: an instance whose checks fail is continued at once by the same
// workgroup with exact divisions (rf_run<.., false>: the factorisation's
// gamma and the quotients by IEEE division) from its last checkpoint (every
// 32 checked steps; else from where it was taken up), so a batch with a few
// such instances neither waits for a second launch nor redoes the steps
// already checked.  Not FAST: only the instances marked for exact
// division (by the throughput form's fast kernel).
template <int NP, int ORDER, bool FAST>
__global__ void __launch_bounds__(64 * rr_waves<NP>()) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_integrate_rf(dh_instance *__restrict__ tab, int ninst, int *__restrict__ counter,
               long long max_steps, long long max_ticks, double *__restrict__ ws_all,
               long long slice)
{
    extern __shared__ double lds[];
    __shared__ int s_next, s_stop;
    const int tid = threadIdx.x;
    double2 *ws = reinterpret_cast<double2 *>(ws_all + (long long)blockIdx.x * slice);
    double *ckp = ws_all + (long long)blockIdx.x * slice + rf_records_doubles<NP>();
    for (;;) {
        __syncthreads();
        if (tid == 0) s_next = atomicAdd(counter, 1);
        __syncthreads();
        const int id = __builtin_amdgcn_readfirstlane(s_next);
        if (id >= ninst) break;
        dh_instance *ip = tab + id;
        const int n = __builtin_amdgcn_readfirstlane(ip->n);
        if (ip->D != 2 || n <= NP - RR_SPAN || n > NP) {
            if (tid == 0) ip->status = -1;
            continue;
        }
        __builtin_assume(n > NP - RR_SPAN && n <= NP);
        if (FAST == ((__builtin_amdgcn_readfirstlane(ip->reserved) & 1) != 0)) continue;
        const long long ld = ip->ld > 0 ? ip->ld : n;
        RfStart at{gbl(ip->phi_dev), ld, ip->cur_seg, ip->cur_step,
                   max_steps > 0 ? max_steps : LLONG_MAX};
        if (!rf_run<NP, ORDER, FAST>(ip, n, at, max_ticks, ws, ckp, lds, &s_stop)) {
            __syncthreads();   // the checkpoint's stores before its loads
            rf_run<NP, ORDER, false>(ip, n, at, max_ticks, ws, ckp, lds, &s_stop);
        }
    }
}

// ---------------------------------------------------------------------------
// Kernel 2: 1D instances, one lane each, lanes pull instances independently.
// Scratch layout (coalesced over lanes): [q][i][slot], q = xInt, df, W0, W1,
// S, gamma.
// ---------------------------------------------------------------------------
template <int ORDER, int DJ = 0>
__global__ void __launch_bounds__(256) k_integrate_1d(dh_instance *__restrict__ tab, int ninst,
                                                      int *__restrict__ counter,
                                                      double *__restrict__ scratch,
                                                      long long nlanes, int nmax,
                                                      long long max_steps)
{
    const long long slot = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= nlanes) return;
    double *T = scratch + slot;
    const long long qs = (long long)nmax * nlanes;  // stride between tables
    double *txi = T, *tdf = T + qs, *tW0 = T + 2 * qs, *tW1 = T + 3 * qs, *tS = T + 4 * qs;
    double *tg = T + 5 * qs;
    double *tVI = T + 6 * qs, *tDX = T + 7 * qs;   // delta trick only
    const long long ts = nlanes;
    for (;;) {
        const int id = atomicAdd(counter, 1);
        if (id >= ninst) break;
        dh_instance *ip = tab + id;
        const int n = ip->n;
        const double *xx = gbl(ip->xx_dev);
        for (int i = 0; i < n - 1; ++i) txi[i * ts] = 0.5 * (xx[i + 1] + xx[i]);
        for (int i = 0; i < n; ++i) {
            double v;
            if (i == 0) v = 2. / (xx[1] - xx[0]);
            else if (i == n - 1) v = 2. / (xx[n - 1] - xx[n - 2]);
            else v = 2. / ((xx[i + 1] - xx[i]) + (xx[i] - xx[i - 1]));
            tdf[i * ts] = v;
        }
        if constexpr (DJ)
            for (int i = 0; i < n - 1; ++i) tDX[i * ts] = xx[i + 1] - xx[i];
        const double dx0 = xx[1] - xx[0], dxl = xx[n - 1] - xx[n - 2];
        long long seg = ip->cur_seg, step = ip->cur_step;
        const long long nseg = ip->nseg;
        long long budget = max_steps > 0 ? max_steps : LLONG_MAX;
        double *phi = gbl(ip->phi_dev);
        const bool injon = ip->inject & 1;
        const bool frz = ip->frozen & 1;
        while (seg < nseg && budget > 0) {
            const dh_segment *sp = gbl(ip->seg_dev) + seg;
            const double nu = sp->nu[0], g = sp->gamma[0], h = sp->h[0];
            const double bp = sp->beta_p, br = sp->beta_r, bq = sp->beta_q;
            const int chrom = sp->chrom;
            for (int i = 0; i < n - 1; ++i) {
                const double dxi = xx[i + 1] - xx[i];
                const double vi = d_vfunc(xx[i], nu) * bp * br / bq;
                const double vj = d_vfunc(xx[i + 1], nu) * bp * br / bq;
                tW0[i * ts] = vi / (2 * dxi);
                tW1[i * ts] = vj / (2 * dxi);
                tS[i * ts] = d_sel(txi[i * ts], g, h, chrom);
                if constexpr (DJ) tVI[i * ts] = d_vfunc(txi[i * ts], nu) * bp * br / bq;
            }
            const double sel0 = d_sel(xx[0], g, h, chrom), sel1 = d_sel(xx[n - 1], g, h, chrom);
            // 1D: the line is the whole axis, both boundary terms always apply
            // when M has the right sign (dadi/integration1D.c:50-53).
            const bool bf = sel0 <= 0, bl = sel1 >= 0;
            const double bcf = (0.5 / nu - sel0) * 2. / dx0;
            const double bcl = -(-0.5 / nu - sel1) * 2. / dxl;
            const double dt = sp->dt, inv_dt = 1. / dt;
            long long nst = sp->nsteps - step;
            if (nst > budget) nst = budget;
            if (!frz) {
                for (long long k = 0; k < nst; ++k)
                    solve_line<0, ORDER, DJ>(phi, 1, n, txi, tdf, tW0, tW1, tS, ts, nullptr,
                                             nullptr, dt, inv_dt, bf, bcf, bl, bcl,
                                             injon ? 1 : -1, sp->inject, tg, ts,
                                             DeljTabs{tVI, tDX});
            }
            budget -= nst;
            step += nst;
            if (step >= sp->nsteps) {
                ++seg;
                step = 0;
            }
        }
        ip->cur_seg = seg;
        ip->cur_step = step;
        ip->status = seg >= nseg ? 1 : 0;
    }
}

// ---------------------------------------------------------------------------
// Kernel 2b: 1D, one wave per integration, everything in LDS.
//
// Within a segment the tridiagonal matrix of every step is the same (same
// dt, same parameters), so the Thomas factorisation -- gamma_j and bet_j of
// dadi/tridiag.c:22-29, which do not depend on phi -- is formed once per
// segment with exactly the per-step operations and kept in LDS.  A step
// then only runs the right-hand-side recurrence
//   u_j = (r_j - a_j u_{j-1}) / bet_j,   u_j -= gamma_{j+1} u_{j+1},
// which gives the bits of the full per-step Thomas solve.  The divisions by
// bet_j use its precomputed reciprocal (rcp_form_exact below).  The
// recurrences are serial, so one lane runs them while the wave's other lanes
// form the coefficients and the right-hand sides; the latency of a step is a
// few dependent operations per row instead of two full divisions and a
// global-memory round trip.
// ---------------------------------------------------------------------------

constexpr int W1_ARRAYS = 7;   // P, R, A, B (bet), Y (1/bet), G (gamma), T
constexpr int W1_PF = 8;       // rows loaded ahead in the serial recurrences
constexpr int W1_PAD = 2 * W1_PF;   // doubles of padding before and after each array

// num / bet from y = RN(1/bet): q0 = num*y is within an ulp of the quotient,
// the residual num - bet*q0 is exact with an FMA, and q0 + residual*y rounds
// to the correctly rounded quotient (P. Markstein, IBM J. Res. Dev. 34(1),
// 1990) -- the bits of an IEEE division -- as long as nothing leaves the
// normal range.  With 2^-16 <= |bet| < 2^57 (checked once per segment) that
// holds for num = +0 and for normal num of magnitude in [2^-960, 2^1000];
// this tells whether num is such a value.
__device__ __forceinline__ bool rcp_form_exact(double num)
{
    const double an = __builtin_fabs(num);
    return (an >= 0x1p-960 && an <= 0x1p1000) || __double_as_longlong(num) == 0;
}

// Forward recurrence of one step on one lane from row j0 on
// (dadi/tridiag.c:24-29 with bet and gamma known): u_j = (r_j - a_j u_{j-1})
// / bet_j into P, u_{j0-1} taken from P.  MK: through the known reciprocals
// without range checks (the caller checks afterwards and redoes rows from
// the first one outside the exact range with MK = false).  Rows go in blocks
// of W1_PF, the next block's operands loading while the current one is
// solved; the arrays carry W1_PAD doubles of padding on both sides, so the
// look-ahead reads past the end without bounds arithmetic.
struct W1Rows {
    double r[W1_PF], a[W1_PF], b[W1_PF], y[W1_PF];
};

__device__ __forceinline__ void w1_load(W1Rows &B, const double *r, const double *a,
                                        const double *b, const double *y)
{
#pragma unroll
    for (int q = 0; q < W1_PF; ++q) {
        B.r[q] = r[q];
        B.a[q] = a[q];
        B.b[q] = b[q];
        B.y[q] = y[q];
    }
}

template <bool MK>
__device__ __forceinline__ double w1_quot(double num, double bet, double y)
{
    if constexpr (MK) {
        const double q0 = num * y;
        const double e = __builtin_fma(-bet, q0, num);
        return __builtin_fma(e, y, q0);
    } else {
        return num / bet;
    }
}

template <bool MK>
__device__ __forceinline__ double w1_rows(const W1Rows &B, double u, double *p, int count)
{
#pragma unroll
    for (int q = 0; q < W1_PF; ++q) {
        if (q < count) {
            u = w1_quot<MK>(B.r[q] - B.a[q] * u, B.b[q], B.y[q]);
            p[q] = u;
        }
    }
    return u;
}

template <bool MK>
__device__ __forceinline__ void w1_forward(double *__restrict__ P, const double *__restrict__ R,
                                           const double *__restrict__ A,
                                           const double *__restrict__ B,
                                           const double *__restrict__ Y, int n, int j0)
{
    double u = j0 == 0 ? w1_quot<MK>(R[0], B[0], Y[0])
                       : w1_quot<MK>(R[j0] - A[j0] * P[j0 - 1], B[j0], Y[j0]);
    P[j0] = u;
    const double *r = R + j0 + 1, *a = A + j0 + 1, *b = B + j0 + 1, *y = Y + j0 + 1;
    double *p = P + j0 + 1;
    const int rows = n - 1 - j0, full = rows / W1_PF, rem = rows - full * W1_PF;
    // two register blocks used in turn: one is solved while the other loads
    W1Rows X, Z;
    w1_load(X, r, a, b, y);
    int k = 0;
    for (; k + 2 <= full; k += 2) {
        w1_load(Z, r + W1_PF, a + W1_PF, b + W1_PF, y + W1_PF);
        u = w1_rows<MK>(X, u, p, W1_PF);
        w1_load(X, r + 2 * W1_PF, a + 2 * W1_PF, b + 2 * W1_PF, y + 2 * W1_PF);
        u = w1_rows<MK>(Z, u, p + W1_PF, W1_PF);
        r += 2 * W1_PF;
        a += 2 * W1_PF;
        b += 2 * W1_PF;
        y += 2 * W1_PF;
        p += 2 * W1_PF;
    }
    if (k < full) {
        w1_load(Z, r + W1_PF, a + W1_PF, b + W1_PF, y + W1_PF);
        u = w1_rows<MK>(X, u, p, W1_PF);
        w1_rows<MK>(Z, u, p + W1_PF, rem);
    } else {
        w1_rows<MK>(X, u, p, rem);
    }
}

// Back substitution of one step on one lane (dadi/tridiag.c:31-33), blocks
// of W1_PF rows going down, padded arrays as in w1_forward.
__device__ __forceinline__ double w1_back_rows(const double (&cp)[W1_PF],
                                               const double (&cg)[W1_PF], double u, double *p,
                                               int count)
{
#pragma unroll
    for (int q = 0; q < W1_PF; ++q) {
        if (q < count) {
            u = cp[q] - cg[q] * u;
            p[-q] = u;
        }
    }
    return u;
}

__device__ __forceinline__ void w1_back_load(double (&cp)[W1_PF], double (&cg)[W1_PF],
                                             const double *p, const double *g)
{
#pragma unroll
    for (int q = 0; q < W1_PF; ++q) {
        cp[q] = p[-q];
        cg[q] = g[-q];
    }
}

__device__ __forceinline__ void w1_backward(double *__restrict__ P,
                                            const double *__restrict__ G, int n)
{
    double u = P[n - 1];
    double *p = P + n - 2;
    const double *g = G + n - 1;
    const int rows = n - 1, full = rows / W1_PF, rem = rows - full * W1_PF;
    double xp[W1_PF], xg[W1_PF], zp[W1_PF], zg[W1_PF];
    w1_back_load(xp, xg, p, g);
    int k = 0;
    for (; k + 2 <= full; k += 2) {
        w1_back_load(zp, zg, p - W1_PF, g - W1_PF);
        u = w1_back_rows(xp, xg, u, p, W1_PF);
        w1_back_load(xp, xg, p - 2 * W1_PF, g - 2 * W1_PF);
        u = w1_back_rows(zp, zg, u, p - W1_PF, W1_PF);
        p -= 2 * W1_PF;
        g -= 2 * W1_PF;
    }
    if (k < full) {
        w1_back_load(zp, zg, p - W1_PF, g - W1_PF);
        u = w1_back_rows(xp, xg, u, p, W1_PF);
        w1_back_rows(zp, zg, u, p - W1_PF, rem);
    } else {
        w1_back_rows(xp, xg, u, p, rem);
    }
}

template <int ORDER, int DJ>
__global__ void __launch_bounds__(64) k_integrate_1w(dh_instance *__restrict__ tab, int ninst,
                                                     int *__restrict__ counter, int nmax,
                                                     long long max_steps, long long max_ticks)
{
    extern __shared__ double lds[];
    __shared__ int s_id, s_first;
    const int lane = threadIdx.x;
    const int st = nmax + 2 * W1_PAD;   // array stride, padding on both sides
    double *P = lds + W1_PAD, *R = P + st, *A = R + st, *B = A + st, *Y = B + st;
    double *G = Y + st, *T = G + st;
    for (;;) {
        if (lane == 0) s_id = atomicAdd(counter, 1);
        __syncthreads();
        const int id = s_id;
        __syncthreads();
        if (id >= ninst) break;
        dh_instance *ip = tab + id;
        const int n = ip->n;
        const double *xx = gbl(ip->xx_dev);
        double *phi = gbl(ip->phi_dev);
        for (int j = lane; j < n; j += 64) P[j] = phi[j];
        long long seg = ip->cur_seg, step = ip->cur_step;
        const long long nseg = ip->nseg;
        long long budget = max_steps > 0 ? max_steps : LLONG_MAX;
        const bool injon = ip->inject & 1;
        const bool frz = ip->frozen & 1;
        const double dx0 = xx[1] - xx[0], dxl = xx[n - 1] - xx[n - 2];
        __syncthreads();
        const long long t0 = max_ticks > 0 ? (long long)wall_clock64() : 0;
        bool out_of_time = false;
        while (seg < nseg && budget > 0 && !out_of_time) {
            const dh_segment *sp = gbl(ip->seg_dev) + seg;
            long long nst = sp->nsteps - step;
            if (nst > budget) nst = budget;
            if (!frz && nst > 0) {
                const double nu = sp->nu[0], g = sp->gamma[0], h = sp->h[0];
                const double bp = sp->beta_p, br = sp->beta_r, bq = sp->beta_q;
                const int chrom = sp->chrom;
                const double dt = sp->dt, inv_dt = 1. / dt, inj = sp->inject;
                const double sel0 = d_sel(xx[0], g, h, chrom), sel1 = d_sel(xx[n - 1], g, h, chrom);
                // both boundary terms of the 1D line (dadi/integration1D.c:50-53)
                const bool bf = sel0 <= 0, bl = sel1 >= 0;
                const double bcf = (0.5 / nu - sel0) * 2. / dx0;
                const double bcl = -(-0.5 / nu - sel1) * 2. / dxl;
                // interface terms atemp/ctemp (compute_abc_nobc,
                // dadi/integration_shared.c:72-93): NA in Y, NC in T
                for (int i = lane; i < n - 1; i += 64) {
                    const double dxi = xx[i + 1] - xx[i];
                    const double vi = d_vfunc(xx[i], nu) * bp * br / bq;
                    const double vj = d_vfunc(xx[i + 1], nu) * bp * br / bq;
                    const double w0 = vi / (2 * dxi), w1 = vj / (2 * dxi);
                    const double xi = 0.5 * (xx[i + 1] + xx[i]);
                    const double M = d_sel(xi, g, h, chrom);
                    double dl = 0.5;
                    if constexpr (DJ != 0)
                        dl = d_delj<ORDER>(M, d_vfunc(xi, nu) * bp * br / bq, xx[i + 1] - xx[i]);
                    Y[i] = M * dl + w0;
                    T[i] = -M * (1 - dl) + w1;
                }
                __syncthreads();
                // rows: a_j in A, b_j in B, c_{j-1} in G
                for (int j = lane; j < n; j += 64) {
                    const double dfj = j == 0       ? 2. / (xx[1] - xx[0])
                                       : j == n - 1 ? 2. / (xx[n - 1] - xx[n - 2])
                                                    : 2. / ((xx[j + 1] - xx[j]) + (xx[j] - xx[j - 1]));
                    double a = 0., b, cm = 0.;
                    if (j == 0) {
                        const double pa = Y[0];
                        if constexpr (ORDER == DH_ORDER_TV) {
                            b = inv_dt + dfj * pa;
                            if (bf) b += bcf;
                        } else {
                            b = dfj * pa;
                            if (bf) b += bcf;
                            b = b + inv_dt;
                        }
                    } else {
                        const double dfm = j - 1 == 0 ? 2. / (xx[1] - xx[0])
                                                      : 2. / ((xx[j] - xx[j - 1]) +
                                                              (xx[j - 1] - xx[j - 2]));
                        const double pa = Y[j - 1], pc = T[j - 1];
                        a = -dfj * pa;
                        cm = -dfm * pc;
                        if (j < n - 1) {
                            const double na = Y[j];
                            if constexpr (ORDER == DH_ORDER_TV)
                                b = (inv_dt + dfj * pc) + dfj * na;
                            else
                                b = (dfj * na + dfj * pc) + inv_dt;
                        } else {
                            if constexpr (ORDER == DH_ORDER_TV) {
                                b = inv_dt + dfj * pc;
                                if (bl) b += bcl;
                            } else {
                                b = dfj * pc;
                                if (bl) b += bcl;
                                b = b + inv_dt;
                            }
                        }
                    }
                    A[j] = a;
                    B[j] = b;
                    G[j] = cm;
                }
                __syncthreads();
                // factorisation (dadi/tridiag.c:22-29): gamma_j, bet_j
                if (lane == 0) {
                    double bet = B[0];
                    for (int j = 1; j < n; ++j) {
                        const double gj = G[j] / bet;
                        bet = B[j] - A[j] * gj;
                        G[j] = gj;
                        B[j] = bet;
                    }
                }
                __syncthreads();
                bool safe = true;
                for (int j = lane; j < n; j += 64) {
                    const double bj = B[j];
                    Y[j] = 1. / bj;
                    const double ab = __builtin_fabs(bj);
                    safe = safe && ab >= 0x1p-16 && ab < 0x1p57;
                }
                safe = __all(safe);
                __syncthreads();
                const int inj_row = injon ? 1 : -1;
                for (long long k = 0; k < nst; ++k) {
                    // right-hand side (the injected mass first)
                    for (int j = lane; j < n; j += 64) {
                        double pj = P[j];
                        if (j == inj_row) pj = pj + inj;
                        R[j] = (ORDER == DH_ORDER_PRE) ? inv_dt * pj : pj / dt;
                    }
                    __syncthreads();
                    if (safe) {
                        if (lane == 0) w1_forward<true>(P, R, A, B, Y, n, 0);
                        if (lane == 0) s_first = n;
                        __syncthreads();
                        // the first row whose numerator was outside the range
                        // where the reciprocal form is exact (the numerators
                        // are re-formed with the same operations)
                        for (int j = lane; j < n; j += 64) {
                            const double num = j == 0 ? R[0] : R[j] - A[j] * P[j - 1];
                            if (!rcp_form_exact(num)) {
                                atomicMin(&s_first, j);
                                break;
                            }
                        }
                        __syncthreads();
                        const int jb = s_first;
                        if (lane == 0 && jb < n) w1_forward<false>(P, R, A, B, Y, n, jb);
                    } else if (lane == 0) {
                        w1_forward<false>(P, R, A, B, Y, n, 0);
                    }
                    if (lane == 0) w1_backward(P, G, n);
                    __syncthreads();
                    // the clock is a scalar read: the wave agrees on it
                    if (max_ticks > 0 && (long long)wall_clock64() - t0 >= max_ticks) {
                        nst = k + 1;
                        out_of_time = true;
                        break;
                    }
                }
            }
            budget -= nst;
            step += nst;
            if (step >= sp->nsteps) {
                ++seg;
                step = 0;
            }
        }
        for (int j = lane; j < n; j += 64) phi[j] = P[j];
        if (lane == 0) {
            ip->cur_seg = seg;
            ip->cur_step = step;
            ip->status = seg >= nseg ? 1 : 0;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Kernel 1c: latency path for 2-5 populations -- one workgroup per instance,
// the Thomas factorisation of every line formed once per segment.
//
// As in the wave-1D kernel, every step of a segment has the same tridiagonal
// matrices, so a_j, bet_j, 1/bet_j and gamma_j of every line of every axis
// are formed once per segment with the per-step operations (row0, row_coeffs
// and dadi/tridiag.c:22-29) into a global workspace laid out [row][line], so
// that the lanes of a wave (consecutive lines) read consecutive addresses.  A
// step then runs only the right-hand-side recurrences, a few dependent
// operations per row instead of the coefficient build and three divisions.
// The forward pass writes u' to the workspace (phi stays intact until the
// back substitution overwrites it), so that a line whose numerators leave the
// range where the reciprocal form is exact can be redone with divisions from
// its first such row.  More memory traffic per point than k_integrate_nd
// (the factorisation is read every step): for few instances, where latency
// rather than bandwidth bounds a step.
// ---------------------------------------------------------------------------

#ifndef DH_FZ_PF
#define DH_FZ_PF 8
#endif
constexpr int FZ_PF = DH_FZ_PF;   // rows per block (two blocks in flight)

// Workspace doubles per workgroup: a, bet, 1/bet, gamma per axis, plus u'.
long long fz_slice(int D, int n)
{
    long long N = 1;
    for (int q = 0; q < D; ++q) N *= n;
    return (4LL * D + 1) * N;
}

__device__ __forceinline__ bool fz_bet_ok(double bet)
{
    const double ab = __builtin_fabs(bet);
    return ab >= 0x1p-16 && ab < 0x1p57;
}

// Factorisation of every line of axis `a`: lines over the lanes of the
// workgroup, tile t of 64 lines on wave t % nw (the same mapping as the
// per-step sweeps, so each lane reads back only what it wrote).
template <int D, int ORDER>
__device__ void fz_factor_axis(int a, int n, long long L, const dh_segment *sp,
                               const double *sx, const double *sxi, const double *sdf,
                               const double *sdx, const double *W0, const double *W1,
                               const double *S, double dt, double inv_dt, double *Fa,
                               long long tile0, long long tstep, int lane, bool &safe)
{
    constexpr int NO = D - 1;
    const long long N = L * n;
    double *A = Fa, *Bt = Fa + N, *Yr = Fa + 2 * N, *G = Fa + 3 * N;
    double mig[NO];
#pragma unroll
    for (int q = 0; q < NO; ++q) mig[q] = sp->mig[a][q];
    const double nu = sp->nu[a], ga = sp->gamma[a], ha = sp->h[a];
    for (long long t = tile0; t * 64 < L; t += tstep) {
        const long long line = t * 64 + lane;
        if (line >= L) continue;
        LineSetup<D, ORDER> Ls;
        line_setup<D, ORDER>(Ls, line, n, a, sx, mig, nu, ga, ha, sdx[0], sdx[n - 2], false, 0,
                             sp->chrom);
        double bet, u0, pa, pc;
        row0<NO, ORDER>(0.0, sxi, sdf, W0, W1, S, Ls.c, mig, dt, inv_dt, Ls.bcf_on, Ls.bcf, bet,
                        u0, pa, pc);
        A[line] = 0.;
        Bt[line] = bet;
        Yr[line] = 1. / bet;
        G[line] = 0.;
        safe = safe && fz_bet_ok(bet);
        for (int j = 1; j < n; ++j) {
            double aj, bj, cm;
            row_coeffs<NO, ORDER>(j, n, sxi, sdf, W0, W1, S, 1, Ls.c, mig, inv_dt, Ls.bcl_on,
                                  Ls.bcl, pa, pc, aj, bj, cm);
            const double g = cm / bet;
            bet = bj - aj * g;
            const long long o = (long long)j * L + line;
            A[o] = aj;
            Bt[o] = bet;
            Yr[o] = 1. / bet;
            G[o] = g;
            safe = safe && fz_bet_ok(bet);
        }
    }
}

template <int ORDER>
__device__ __forceinline__ double fz_rhs(double p, double dt, double inv_dt)
{
    return (ORDER == DH_ORDER_PRE) ? inv_dt * p : p / dt;
}

// Operands of FZ_PF rows of one line: phi, a, bet, 1/bet.  Rows past the
// end of the line repeat its last row (only the look-ahead reads them).
struct FzRows {
    double p[FZ_PF], a[FZ_PF], b[FZ_PF], y[FZ_PF];
};

__device__ __forceinline__ void fz_load(FzRows &R, const double *p, long long s,
                                        const double *A, const double *Bt, const double *Yr,
                                        long long L, int j0, int n)
{
#pragma unroll
    for (int q = 0; q < FZ_PF; ++q) {
        const int jj = j0 + q < n ? j0 + q : n - 1;
        const long long o = (long long)jj * L;
        R.p[q] = p[jj * s];
        R.a[q] = A[o];
        R.b[q] = Bt[o];
        R.y[q] = Yr[o];
    }
}

template <int ORDER, bool MK>
__device__ __forceinline__ double fz_rows(const FzRows &R, double u, double *U, long long L,
                                          int j0, int n, double dt, double inv_dt, int &bad)
{
#pragma unroll
    for (int q = 0; q < FZ_PF; ++q) {
        const int j = j0 + q;
        if (j < n) {
            const double num = fz_rhs<ORDER>(R.p[q], dt, inv_dt) - R.a[q] * u;
            if constexpr (MK) {
                const double q0 = num * R.y[q];
                u = __builtin_fma(__builtin_fma(-R.b[q], q0, num), R.y[q], q0);
                if (bad == n && !rcp_form_exact(num)) bad = j;
            } else {
                u = num / R.b[q];
            }
            U[(long long)j * L] = u;
        }
    }
    return u;
}

__device__ __forceinline__ void fz_back_load(double (&cu)[FZ_PF], double (&cg)[FZ_PF],
                                             const double *U, const double *G, long long L,
                                             int j0)
{
#pragma unroll
    for (int q = 0; q < FZ_PF; ++q) {
        const int jj = j0 - q >= 0 ? j0 - q : 0;
        cu[q] = U[(long long)jj * L];
        cg[q] = G[(long long)(jj + 1) * L];
    }
}

__device__ __forceinline__ double fz_back_rows(const double (&cu)[FZ_PF],
                                               const double (&cg)[FZ_PF], double u, double *p,
                                               long long s, int j0)
{
#pragma unroll
    for (int q = 0; q < FZ_PF; ++q) {
        const int j = j0 - q;
        if (j >= 0) {
            u = cu[q] - cg[q] * u;
            p[j * s] = u;
        }
    }
    return u;
}

// One step's sweep of axis `a` from the factorisation.  MK: divisions through
// the known reciprocals (each line redone with divisions from its first row
// outside the exact range).  Rows go in blocks of FZ_PF from two register
// blocks used in turn, one loading while the other is solved.
template <int D, int ORDER, bool MK>
__device__ void fz_sweep_axis(double *__restrict__ phi, int n, long long ld, long long L, int a,
                              bool first_axis, int injm, const dh_segment *sp, const double *sx,
                              double dx0, double dxl, double dt, double inv_dt,
                              const double *__restrict__ Fa, double *__restrict__ Uw,
                              long long tile0, long long tstep, int lane)
{
    const long long N = L * n;
    double mig[D - 1];
#pragma unroll
    for (int q = 0; q < D - 1; ++q) mig[q] = sp->mig[a][q];
    const double inj = sp->inject;
    for (long long t = tile0; t * 64 < L; t += tstep) {
        const long long line = t * 64 + lane;
        if (line >= L) continue;
        const double *A = Fa + line, *Bt = Fa + N + line, *Yr = Fa + 2 * N + line;
        const double *G = Fa + 3 * N + line;
        double *U = Uw + line;
        LineSetup<D, ORDER> Ls;
        line_setup<D, ORDER>(Ls, line, n, a, sx, mig, sp->nu[a], sp->gamma[a], sp->h[a], dx0,
                             dxl, first_axis, injm, sp->chrom);
        long long off, s;
        line_addr<D>(n, ld, a, line, off, s);
        double *p = phi + off;
        if (Ls.inj_row >= 0) p[Ls.inj_row * s] = p[Ls.inj_row * s] + inj;

        // forward, rows 0..n-1, u' into U
        int bad = n;
        double u;
        {
            const double r0 = fz_rhs<ORDER>(p[0], dt, inv_dt);
            if constexpr (MK) {
                const double q0 = r0 * Yr[0];
                u = __builtin_fma(__builtin_fma(-Bt[0], q0, r0), Yr[0], q0);
                if (!rcp_form_exact(r0)) bad = 0;
            } else {
                u = r0 / Bt[0];
            }
            U[0] = u;
        }
        FzRows X, Z;
        fz_load(X, p, s, A, Bt, Yr, L, 1, n);
        int j0 = 1;
        for (; j0 + FZ_PF < n; j0 += 2 * FZ_PF) {
            fz_load(Z, p, s, A, Bt, Yr, L, j0 + FZ_PF, n);
            u = fz_rows<ORDER, MK>(X, u, U, L, j0, n, dt, inv_dt, bad);
            if (j0 + 2 * FZ_PF < n) fz_load(X, p, s, A, Bt, Yr, L, j0 + 2 * FZ_PF, n);
            u = fz_rows<ORDER, MK>(Z, u, U, L, j0 + FZ_PF, n, dt, inv_dt, bad);
        }
        if (j0 < n) u = fz_rows<ORDER, MK>(X, u, U, L, j0, n, dt, inv_dt, bad);
        if (MK && bad < n) {
            // rows from the first one outside the exact range, with divisions
            u = bad == 0 ? 0. : U[(long long)(bad - 1) * L];
            for (int j = bad; j < n; ++j) {
                const long long o = (long long)j * L;
                const double r = fz_rhs<ORDER>(p[j * s], dt, inv_dt);
                u = (j == 0 ? r : r - A[o] * u) / Bt[o];
                U[o] = u;
            }
        }

        // back substitution (dadi/tridiag.c:31-33), u' from U, result into phi
        u = U[(long long)(n - 1) * L];
        p[(long long)(n - 1) * s] = u;
        double xu[FZ_PF], xg[FZ_PF], zu[FZ_PF], zg[FZ_PF];
        fz_back_load(xu, xg, U, G, L, n - 2);
        int k0 = n - 2;
        for (; k0 - FZ_PF >= 0; k0 -= 2 * FZ_PF) {
            fz_back_load(zu, zg, U, G, L, k0 - FZ_PF);
            u = fz_back_rows(xu, xg, u, p, s, k0);
            if (k0 - 2 * FZ_PF >= 0) fz_back_load(xu, xg, U, G, L, k0 - 2 * FZ_PF);
            u = fz_back_rows(zu, zg, u, p, s, k0 - FZ_PF);
        }
        if (k0 >= 0) u = fz_back_rows(xu, xg, u, p, s, k0);
    }
}

template <int D, int ORDER>
__global__ void __launch_bounds__(512) k_integrate_fz(dh_instance *__restrict__ tab, int ninst,
                                                       int *__restrict__ counter,
                                                       long long max_steps,
                                                       double *__restrict__ ws, long long slice,
                                                       long long max_ticks)
{
    extern __shared__ double lds[];
    __shared__ int s_next, s_stop, s_unsafe;
    const int nthr = blockDim.x, tid = threadIdx.x;
    const int nw = nthr >> 6, wave = tid >> 6, lane = tid & 63;
    double *F = ws + (long long)blockIdx.x * slice;
    for (;;) {
        __syncthreads();
        if (tid == 0) s_next = atomicAdd(counter, 1);
        __syncthreads();
        const int id = __builtin_amdgcn_readfirstlane(s_next);
        if (id >= ninst) break;
        dh_instance *ip = tab + id;
        const int n = ip->n;
        long long L = 1;
        for (int q = 0; q < D - 1; ++q) L *= n;
        const long long N = L * n;
        double *U = F + 4LL * D * N;
        double *sx = lds, *sxi = sx + n, *sdf = sxi + n, *sdx = sdf + n;
        double *tabs = sdx + n;
        build_grid_tables(gbl(ip->xx_dev), n, sx, sxi, sdf, sdx, tid, nthr);
        long long seg = ip->cur_seg, step = ip->cur_step;
        const long long nseg = ip->nseg;
        long long budget = max_steps > 0 ? max_steps : LLONG_MAX;
        const int frozen = ip->frozen, injm = ip->inject;
        double *phi = gbl(ip->phi_dev);
        const long long ld = ip->ld > 0 ? ip->ld : n;
        const long long t0 = max_ticks > 0 ? (long long)wall_clock64() : 0;
        bool out_of_time = false;
        while (seg < nseg && budget > 0 && !out_of_time) {
            const dh_segment *sp = gbl(ip->seg_dev) + seg;
            long long nst = sp->nsteps - step;
            if (nst > budget) nst = budget;
            __syncthreads();
            for (int a = 0; a < D; ++a) {
                if ((frozen >> a) & 1) continue;
                double *W0 = tabs + LDS_TABLES_PER_AXIS * a * n;
                build_axis_tables<D>(sp, a, n, sx, sxi, sdx, W0, W0 + n, W0 + 2 * n, tid, nthr);
            }
            if (tid == 0) s_unsafe = 0;
            __syncthreads();
            const double dt = sp->dt, inv_dt = 1. / dt;
            bool safe = true;
            for (int a = 0; a < D; ++a) {
                if ((frozen >> a) & 1) continue;
                const double *W0 = tabs + LDS_TABLES_PER_AXIS * a * n;
                fz_factor_axis<D, ORDER>(a, n, L, sp, sx, sxi, sdf, sdx, W0, W0 + n, W0 + 2 * n,
                                         dt, inv_dt, F + 4LL * a * N, wave, nw, lane, safe);
            }
            if (!safe) s_unsafe = 1;
            __syncthreads();
            const bool mk = s_unsafe == 0;
            for (long long k = 0; k < nst; ++k) {
                bool first = true;
                for (int a = 0; a < D; ++a) {
                    if ((frozen >> a) & 1) continue;
                    if (mk)
                        fz_sweep_axis<D, ORDER, true>(phi, n, ld, L, a, first, injm, sp, sx,
                                                      sdx[0], sdx[n - 2], dt, inv_dt,
                                                      F + 4LL * a * N, U, wave, nw, lane);
                    else
                        fz_sweep_axis<D, ORDER, false>(phi, n, ld, L, a, first, injm, sp, sx,
                                                       sdx[0], sdx[n - 2], dt, inv_dt,
                                                       F + 4LL * a * N, U, wave, nw, lane);
                    __syncthreads();
                    first = false;
                }
                if (max_ticks > 0) {
                    if (tid == 0) s_stop = (long long)wall_clock64() - t0 >= max_ticks;
                    __syncthreads();
                    if (s_stop) {
                        nst = k + 1;
                        out_of_time = true;
                        break;
                    }
                }
            }
            budget -= nst;
            step += nst;
            if (step >= sp->nsteps) {
                ++seg;
                step = 0;
            }
        }
        if (tid == 0) {
            ip->cur_seg = seg;
            ip->cur_step = step;
            ip->status = seg >= nseg ? 1 : 0;
        }
    }
}

// ---------------------------------------------------------------------------
// Kernel 3: one axis of one step of one instance over many workgroups.
// ---------------------------------------------------------------------------
template <int D, int ORDER>
__global__ void __launch_bounds__(256) k_sweep_axis(dh_instance inst, int a, int first_axis,
                                                    const long long *__restrict__ base,
                                                    long long off,
                                                    const int *__restrict__ seg_of_step,
                                                    long long total)
{
    const long long st = *base + off;
    if (st >= total) return;
    extern __shared__ double lds[];
    const int n = inst.n;
    const int nthr = blockDim.x, tid = threadIdx.x;
    const int nw = nthr >> 6, wave = tid >> 6, lane = tid & 63;
    const dh_segment *sp = gbl(inst.seg_dev) + seg_of_step[st];
    double *sx = lds, *sxi = sx + n, *sdf = sxi + n, *sdx = sdf + n;
    double *W0 = sdx + n, *W1 = W0 + n, *S = W1 + n;
    double *gamcol = S + n + (long long)wave * 64 * n + lane;
    build_grid_tables(gbl(inst.xx_dev), n, sx, sxi, sdf, sdx, tid, nthr);
    __syncthreads();
    build_axis_tables<D>(sp, a, n, sx, sxi, sdx, W0, W1, S, tid, nthr);
    __syncthreads();
    const double dt = sp->dt, inv_dt = 1. / dt;
    sweep_tiles<D, ORDER, 0, 0, 0, 0, SWEEP_PF, true>(gbl(inst.phi_dev), n, inst.ld > 0 ? inst.ld : n, a,
                                   first_axis != 0, inst.inject, sp, sx, sxi, sdf, sdx, W0, W1,
                                   S, dt, inv_dt, gamcol, nullptr,
                                   (long long)blockIdx.x * nw + wave,
                                   (long long)gridDim.x * nw, lane);
}

// ---------------------------------------------------------------------------
// Kernel 3b: one axis of one step of a batch of integrations, each one's
// lines spread over several workgroups.  CK = 0: each lane's gamma column in
// LDS (as k_sweep_axis); CK = 8: checkpointed Thomas, for grids whose
// columns do not fit.
// ---------------------------------------------------------------------------
template <int D, int ORDER, int CK>
__global__ void __launch_bounds__(256) k_sweep_batch(const dh_sweep_item *__restrict__ items,
                                                     const int *__restrict__ wg_item, int a,
                                                     const long long *__restrict__ base,
                                                     long long off)
{
    const dh_sweep_item &it = items[wg_item[blockIdx.x]];
    const long long st = *base + off + it.step_shift;
    if (st >= it.total_steps) return;
    const dh_instance &I = it.inst;
    if ((I.frozen >> a) & 1) return;
    extern __shared__ double lds[];
    const int n = I.n;
    const int nthr = blockDim.x, tid = threadIdx.x;
    const int nw = nthr >> 6, wave = tid >> 6, lane = tid & 63;
    const int col = CK ? (n - 1) / (CK ? CK : 1) + 1 : n;
    const dh_segment *sp = gbl(I.seg_dev) + gbl(it.seg_of_step_dev)[st];
    double *sx = lds, *sxi = sx + n, *sdf = sxi + n, *sdx = sdf + n;
    double *W0 = sdx + n, *W1 = W0 + n, *S = W1 + n;
    build_grid_tables(gbl(I.xx_dev), n, sx, sxi, sdf, sdx, tid, nthr);
    __syncthreads();
    build_axis_tables<D>(sp, a, n, sx, sxi, sdx, W0, W1, S, tid, nthr);
    __syncthreads();
    const double dt = sp->dt, inv_dt = 1. / dt;
    double *gamcol = S + n + (long long)wave * 64 * col + lane;
    double *stage = S + n + (long long)nw * 64 * col + (long long)wave * 64 * (STAGE_C + 1);
    sweep_tiles<D, ORDER, CK, 0, 0, 0, SWEEP_PF, true>(gbl(I.phi_dev), n, I.ld > 0 ? I.ld : n, a,
                                    a == it.first_axis, I.inject, sp, sx, sxi, sdf, sdx, W0, W1,
                                    S, dt, inv_dt, gamcol, stage,
                                    (long long)(blockIdx.x - it.wg_begin) * nw + wave,
                                    (long long)it.wgs * nw, lane);
}

__global__ void k_advance(long long *base, long long by)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *base += by;
}

__global__ void k_set_int(int *p, int v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ void k_set_ll(long long *p, long long v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// ---------------------------------------------------------------------------
// Sampling: s = dphi/dx and c1 = (phi - s x)/(m+1) per interval, stored
// interval-major so the contraction reads them coalesced.
// ---------------------------------------------------------------------------
__global__ void k_sample_prep(const double *__restrict__ in, long long P, int n,
                              const double *__restrict__ xx, int m, double *__restrict__ sT,
                              double *__restrict__ cT)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long tot = P * (n - 1);
    if (e >= tot) return;
    const long long i = e / P, p = e - i * P;
    const double f0 = in[p * n + i], f1 = in[p * n + i + 1];
    const double s = (f1 - f0) / (xx[i + 1] - xx[i]);   // dadi/Spectrum_mod.py:1527
    const double c1 = (f0 - s * xx[i]) / (m + 1);        // dadi/Spectrum_mod.py:1528
    sT[e] = s;
    cT[e] = c1;
}

// Error-free transformations (Knuth's TwoSum; TwoProduct through the fused
// multiply-add, which is exact): a + b = s + e and a * b = p + e exactly.
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e)
{
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ void two_prod(double a, double b, double &p, double &e)
{
    p = a * b;
    e = fma(a, b, -p);
}

// out[k][p] = sum_i c1[p][i] db1[k][i] + fac_k sum_i s[p][i] db2[k][i]
// (dadi/Spectrum_mod.py:1527-1532, fac_k = (k+1)/((m+1)(m+2))).  The sums are
// compensated (Ogita, Rump and Oishi's Dot2, SIAM J. Sci. Comput. 26, 2005):
// each product and partial sum carries its exact rounding error, so the result
// is the exactly rounded value of the formula up to an error of order
// eps^2 times the sum of the magnitudes of the terms.  Entries next to the
// spectrum's corners cancel strongly (terms ~1e4 times the result), where an
// uncompensated sum in any order is off by ~1e-12 relative before
// extrapolation amplifies it.
__global__ void k_sample_contract(const double *__restrict__ sT, const double *__restrict__ cT,
                                  long long P, int n, int m, const double *__restrict__ db1,
                                  const double *__restrict__ db2, double *__restrict__ out)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)(m + 1) * P) return;
    const long long k = e / P, p = e - k * P;
    const double *d1 = db1 + k * (n - 1), *d2 = db2 + k * (n - 1);
    double s1 = 0., c1 = 0., s2 = 0., c2 = 0.;
    for (int i = 0; i < n - 1; ++i) {
        double q, eq, es;
        two_prod(cT[(long long)i * P + p], d1[i], q, eq);
        two_sum(s1, q, s1, es);
        c1 = c1 + (es + eq);
        two_prod(sT[(long long)i * P + p], d2[i], q, eq);
        two_sum(s2, q, s2, es);
        c2 = c2 + (es + eq);
    }
    const double fac = (double)(k + 1) / (double)((m + 1) * (m + 2));
    double q, eq, t, et;
    two_prod(s2, fac, q, eq);
    two_sum(s1, q, t, et);
    out[e] = t + (et + ((c1 + eq) + c2 * fac));
}

// Direct (trapezoid) sampling: out[k][p] = sum_i A[k][i] * in[p][i], where
// A[k][i] holds the binomial sampling factor of entry k at grid point i times
// the trapezoid weight of that point (dadi/Spectrum_mod.py:1263-1279,
// 1350-1380, 1550-1592, 1594-1643).
__global__ void k_contract_axis(const double *__restrict__ in, long long P, int n,
                                const double *__restrict__ A, int m1, double *__restrict__ out)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)m1 * P) return;
    const long long k = e / P, p = e - k * P;
    const double *a = A + k * n, *row = in + p * n;
    double t = 0.;
    for (int i = 0; i < n; ++i) t = t + a[i] * row[i];
    out[e] = t;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
namespace {

template <int D, int ORDER>
const void *nd_kernel(int chunk, int recompute, int gg, int dj)
{
    if (dj) return reinterpret_cast<const void *>(&k_integrate_nd<D, ORDER, 0, 0, 0, 1>);
    if (gg) return reinterpret_cast<const void *>(&k_integrate_nd<D, ORDER, 0, 0, 1>);
    if (chunk == 8 && recompute)
        return reinterpret_cast<const void *>(&k_integrate_nd<D, ORDER, 8, 1, 0>);
    if (chunk == 8) return reinterpret_cast<const void *>(&k_integrate_nd<D, ORDER, 8, 0, 0>);
    return reinterpret_cast<const void *>(&k_integrate_nd<D, ORDER, 0, 0, 0>);
}

const void *pick_nd(int D, int order, int chunk, int recompute, int gg, int dj)
{
    switch (D * 4 + order) {
#define DH_CASE(d, o) case d * 4 + o: return nd_kernel<d, o>(chunk, recompute, gg, dj);
        DH_CASE(1, 0) DH_CASE(1, 1) DH_CASE(1, 2)
        DH_CASE(2, 0) DH_CASE(2, 1) DH_CASE(2, 2)
        DH_CASE(3, 0) DH_CASE(3, 1) DH_CASE(3, 2)
        DH_CASE(4, 0) DH_CASE(4, 1) DH_CASE(4, 2)
        DH_CASE(5, 0) DH_CASE(5, 1) DH_CASE(5, 2)
#undef DH_CASE
    }
    return nullptr;
}

template <int D, int ORDER>
const void *pk_kernel() { return reinterpret_cast<const void *>(&k_integrate_packed<D, ORDER>); }

const void *pick_packed(int D, int order)
{
    switch (D * 4 + order) {
#define DH_CASE(d, o) case d * 4 + o: return pk_kernel<d, o>();
        DH_CASE(2, 0) DH_CASE(2, 1) DH_CASE(2, 2)
        DH_CASE(3, 0) DH_CASE(3, 1) DH_CASE(3, 2)
        DH_CASE(4, 0) DH_CASE(4, 1) DH_CASE(4, 2)
        DH_CASE(5, 0) DH_CASE(5, 1) DH_CASE(5, 2)
#undef DH_CASE
    }
    return nullptr;
}

// Statically allocated LDS of the packed kernel (slot table and flag).
constexpr size_t PACK_STATIC_LDS = sizeof(PackSlot) * DH_PACK_MAX + 16;

// Dynamic LDS of the packed kernel.
size_t packed_dyn_lds(int D, int n, int waves, int slots)
{
    const size_t col = (size_t)((n - 1) / 8 + 1);
    return sizeof(double) *
           ((size_t)n * (LDS_GRID_TABLES + LDS_TABLES_PER_AXIS * (size_t)D * (size_t)slots) +
            (size_t)waves * 64 * (col + STAGE_C + 1));
}

template <int D, int ORDER>
const void *sw_kernel() { return reinterpret_cast<const void *>(&k_sweep_axis<D, ORDER>); }

const void *pick_sweep(int D, int order)
{
    switch (D * 4 + order) {
#define DH_CASE(d, o) case d * 4 + o: return sw_kernel<d, o>();
        DH_CASE(1, 0) DH_CASE(1, 1) DH_CASE(1, 2)
        DH_CASE(2, 0) DH_CASE(2, 1) DH_CASE(2, 2)
        DH_CASE(3, 0) DH_CASE(3, 1) DH_CASE(3, 2)
        DH_CASE(4, 0) DH_CASE(4, 1) DH_CASE(4, 2)
        DH_CASE(5, 0) DH_CASE(5, 1) DH_CASE(5, 2)
#undef DH_CASE
    }
    return nullptr;
}

template <int D, int ORDER>
const void *fz_kernel() { return reinterpret_cast<const void *>(&k_integrate_fz<D, ORDER>); }

// register-resident 2D kernel: the row counts NP it is built for (each takes
// n in (NP - 8, NP]); the coefficient orders of 2 populations (TV, PRE); the
// shared-reciprocal (fast) and the exactly dividing form
template <int NP, bool FAST>
const void *rr_kernel(int order)
{
    switch (order) {
    case DH_ORDER_TV:
        return reinterpret_cast<const void *>(&k_integrate_rr<NP, DH_ORDER_TV, FAST>);
    case DH_ORDER_PRE:
        return reinterpret_cast<const void *>(&k_integrate_rr<NP, DH_ORDER_PRE, FAST>);
    }
    return nullptr;
}

template <int NP>
size_t rr_lds_bytes_of() { return sizeof(double) * rr_lds_doubles<NP>(); }

// kernel and LDS bytes for n rows (null if n has no kernel)
const void *pick_rr(int n, int order, bool fast, size_t *lds, int *waves)
{
#define DH_RR(np)                                                                   \
    if (n > (np) - RR_SPAN && n <= (np)) {                                          \
        *lds = rr_lds_bytes_of<np>();                                               \
        *waves = rr_waves<np>();                                                    \
        return fast ? rr_kernel<np, true>(order) : rr_kernel<np, false>(order);     \
    }
    DH_RR(40) DH_RR(48) DH_RR(56) DH_RR(64) DH_RR(72) DH_RR(80) DH_RR(88) DH_RR(96) DH_RR(104)
#undef DH_RR
    return nullptr;
}

// latency form of the register-resident kernel (k_integrate_rf): kernel, LDS
// bytes, waves and workspace doubles per workgroup for n rows (null if none)
template <int NP, bool FAST>
const void *rf_kernel(int order)
{
    switch (order) {
    case DH_ORDER_TV:
        return reinterpret_cast<const void *>(&k_integrate_rf<NP, DH_ORDER_TV, FAST>);
    case DH_ORDER_PRE:
        return reinterpret_cast<const void *>(&k_integrate_rf<NP, DH_ORDER_PRE, FAST>);
    }
    return nullptr;
}

const void *pick_rf(int n, int order, bool fast, size_t *lds, int *waves, long long *slice,
                    int *rows = nullptr)
{
#define DH_RF(np)                                                                   \
    if (n > (np) - RR_SPAN && n <= (np)) {                                          \
        *lds = sizeof(double) * rf_lds_doubles<np>();                               \
        *waves = rr_waves<np>();                                                    \
        *slice = rf_slice_doubles<np>();                                            \
        if (rows) *rows = np;                                                       \
        return fast ? rf_kernel<np, true>(order) : rf_kernel<np, false>(order);     \
    }
    DH_RF(40) DH_RF(48) DH_RF(56) DH_RF(64) DH_RF(72) DH_RF(80) DH_RF(88) DH_RF(96) DH_RF(104)
    DH_RF(112) DH_RF(120) DH_RF(128)
#undef DH_RF
    return nullptr;
}

const void *pick_fz(int D, int order)
{
    switch (D * 4 + order) {
#define DH_CASE(d, o) case d * 4 + o: return fz_kernel<d, o>();
        DH_CASE(2, 0) DH_CASE(2, 1) DH_CASE(2, 2)
        DH_CASE(3, 0) DH_CASE(3, 1) DH_CASE(3, 2)
        DH_CASE(4, 0) DH_CASE(4, 1) DH_CASE(4, 2)
        DH_CASE(5, 0) DH_CASE(5, 1) DH_CASE(5, 2)
#undef DH_CASE
    }
    return nullptr;
}

template <int D, int ORDER>
const void *sb_kernel(int ck)
{
    return ck ? reinterpret_cast<const void *>(&k_sweep_batch<D, ORDER, 8>)
              : reinterpret_cast<const void *>(&k_sweep_batch<D, ORDER, 0>);
}

const void *pick_sweep_batch(int D, int order, int ck)
{
    switch (D * 4 + order) {
#define DH_CASE(d, o) case d * 4 + o: return sb_kernel<d, o>(ck);
        DH_CASE(2, 0) DH_CASE(2, 1) DH_CASE(2, 2)
        DH_CASE(3, 0) DH_CASE(3, 1) DH_CASE(3, 2)
        DH_CASE(4, 0) DH_CASE(4, 1) DH_CASE(4, 2)
        DH_CASE(5, 0) DH_CASE(5, 1) DH_CASE(5, 2)
#undef DH_CASE
    }
    return nullptr;
}

const void *pick_1w(int order, int dj)
{
    switch (order * 2 + (dj ? 1 : 0)) {
    case 0: return reinterpret_cast<const void *>(&k_integrate_1w<0, 0>);
    case 1: return reinterpret_cast<const void *>(&k_integrate_1w<0, 1>);
    case 2: return reinterpret_cast<const void *>(&k_integrate_1w<1, 0>);
    case 3: return reinterpret_cast<const void *>(&k_integrate_1w<1, 1>);
    case 4: return reinterpret_cast<const void *>(&k_integrate_1w<2, 0>);
    case 5: return reinterpret_cast<const void *>(&k_integrate_1w<2, 1>);
    }
    return nullptr;
}

const void *pick_1d(int order, int dj)
{
    switch (order * 2 + (dj ? 1 : 0)) {
    case 0: return reinterpret_cast<const void *>(&k_integrate_1d<0, 0>);
    case 1: return reinterpret_cast<const void *>(&k_integrate_1d<0, 1>);
    case 2: return reinterpret_cast<const void *>(&k_integrate_1d<1, 0>);
    case 3: return reinterpret_cast<const void *>(&k_integrate_1d<1, 1>);
    case 4: return reinterpret_cast<const void *>(&k_integrate_1d<2, 0>);
    case 5: return reinterpret_cast<const void *>(&k_integrate_1d<2, 1>);
    }
    return nullptr;
}

size_t nd_lds(int D, int n, int waves, int chunk, int recompute, int gg, int dj = 0)
{
    const size_t col = gg ? 0
                          : chunk ? (size_t)((n - 1) / chunk + 1) * (recompute ? 2 : 1) : (size_t)n;
    const size_t stage = (chunk || gg) ? 64 * (size_t)((recompute ? chunk : STAGE_C) + 1) : 0;
    const size_t tpa = LDS_TABLES_PER_AXIS + (dj ? 1 : 0);
    return sizeof(double) * ((size_t)n * (LDS_GRID_TABLES + tpa * (size_t)D) +
                             (size_t)waves * (64 * col + stage));
}

size_t sweep_lds(int n, int waves)
{
    return sizeof(double) * (size_t)n * (LDS_GRID_TABLES + LDS_TABLES_PER_AXIS + 64 * (size_t)waves);
}

// Persistent workgroups of an nD launch: as many as fit on the device
// (LDS-limited, at most ~16 waves per compute unit), never more than there
// are instances.
int nd_groups(size_t lds, int waves, int groups, int ninst)
{
    if (groups <= 0) {
        int per_cu = lds ? (int)(g_max_lds / lds) : 1;
        if (per_cu < 1) per_cu = 1;
        if (per_cu > 16 / waves + 1) per_cu = 16 / waves + 1;
        groups = g_num_cu * per_cu;
    }
    return groups > ninst ? ninst : groups;
}

int check_init()
{
    if (!g_inited) return fail(DH_EINVAL, "dh_init() has not been called");
    return DH_OK;
}

// cfg->max_us in ticks of wall_clock64(); 0 = no time budget
long long budget_ticks(const dh_launch *cfg)
{
    return cfg->max_us > 0 ? (long long)cfg->max_us * g_wall_khz / 1000 : 0;
}

int set_lds(const void *fn, size_t bytes)
{
    if (bytes > g_max_lds)
        return fail(DH_ERANGE, "LDS request of " + std::to_string(bytes) + " bytes exceeds " +
                                   std::to_string(g_max_lds));
    DH_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return DH_OK;
}

// Cache of instantiated step graphs, keyed by everything captured into them.
// A caller holds a reference to its entry from the lookup until its
// launches are enqueued, so eviction never destroys an executable that is
// about to be launched.  After launching, the caller records an event on its
// stream (one event per stream: launches on a stream complete in order).
// Evicted entries wait in a list and are destroyed, at a later insertion,
// once no caller holds them and every recorded event has completed; nothing
// here waits for the device.
struct GraphEntry {
    hipGraphExec_t exec = nullptr;
    unsigned long long stamp = 0;          // last use, for least-recently-used eviction
    std::mutex mu;                         // guards streams/events
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> events;
    GraphEntry() = default;
    GraphEntry(const GraphEntry &) = delete;
    GraphEntry &operator=(const GraphEntry &) = delete;
    ~GraphEntry()
    {
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        if (exec) (void)hipGraphExecDestroy(exec);
    }
    // completion marker of this caller's launches on st
    hipError_t note_launch(hipStream_t st)
    {
        std::lock_guard<std::mutex> lk(mu);
        for (size_t i = 0; i < streams.size(); ++i)
            if (streams[i] == st) return hipEventRecord(events[i], st);
        hipEvent_t e;
        hipError_t ec = hipEventCreateWithFlags(&e, hipEventDisableTiming);
        if (ec != hipSuccess) return ec;
        streams.push_back(st);
        events.push_back(e);
        return hipEventRecord(e, st);
    }
    // every launch recorded so far has completed (no waiting)
    bool idle()
    {
        std::lock_guard<std::mutex> lk(mu);
        for (hipEvent_t e : events)
            if (hipEventQuery(e) != hipSuccess) return false;
        return true;
    }
};
using GraphRef = std::shared_ptr<GraphEntry>;

constexpr size_t DH_GRAPH_CACHE = 64;
std::mutex g_graph_mu;
std::map<std::string, GraphRef> g_graphs;
std::vector<GraphRef> g_graph_retired;
unsigned long long g_graph_clock = 0;

// Entry for `key`, or null; the returned reference keeps it alive.
GraphRef graph_lookup(const std::string &key)
{
    std::lock_guard<std::mutex> lk(g_graph_mu);
    auto it = g_graphs.find(key);
    if (it == g_graphs.end()) return nullptr;
    it->second->stamp = ++g_graph_clock;
    return it->second;
}

// Adds a freshly instantiated executable (the cache owns it from here on) and
// returns its entry; when another thread inserted the same key meanwhile,
// that entry is returned and `exec` destroyed (it was never launched).
GraphRef graph_insert(const std::string &key, hipGraphExec_t exec)
{
    GraphRef mine = std::make_shared<GraphEntry>();
    mine->exec = exec;
    std::vector<GraphRef> drop;   // destroyed after the lock is released
    GraphRef out;
    {
        std::lock_guard<std::mutex> lk(g_graph_mu);
        // destroy retired entries nobody holds whose launches have completed
        for (size_t i = 0; i < g_graph_retired.size();) {
            GraphRef &r = g_graph_retired[i];
            if (r.use_count() == 1 && r->idle()) {
                drop.push_back(std::move(r));
                r = std::move(g_graph_retired.back());
                g_graph_retired.pop_back();
            } else {
                ++i;
            }
        }
        auto it = g_graphs.find(key);
        if (it != g_graphs.end()) {
            out = it->second;
            drop.push_back(std::move(mine));
        } else {
            while (g_graphs.size() >= DH_GRAPH_CACHE) {
                auto lru = g_graphs.begin();
                for (auto j = g_graphs.begin(); j != g_graphs.end(); ++j)
                    if (j->second->stamp < lru->second->stamp) lru = j;
                g_graph_retired.push_back(std::move(lru->second));
                g_graphs.erase(lru);
            }
            out = mine;
            g_graphs.emplace(key, mine);
        }
        out->stamp = ++g_graph_clock;
    }
    return out;
}

// Scratch of dh_integrate_factored_static (the entry point for callers that
// cannot allocate): the queue counter in the first DH_STATIC_HEAD doubles,
// then the factorisation workspace.  Every device the library runs on holds
// its own copy until the process ends; host code reaches it only through the
// per-device address from hipGetSymbolAddress, never through the name.
constexpr long long DH_STATIC_HEAD = 32;                       // 256 bytes
constexpr long long DH_STATIC_WORKSPACE = 4LL << 20;           // doubles (32 MiB)
constexpr int DH_MAX_DEVICES = 64;
double *g_scratch_dev[DH_MAX_DEVICES] = {};
std::mutex g_scratch_mu;

}  // namespace

static __device__ __align__(256) double scratch[DH_STATIC_HEAD + DH_STATIC_WORKSPACE];

namespace {

// Address of `scratch` on the current device, resolved once per device.
int static_scratch(double **out)
{
    int dev = 0;
    DH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= DH_MAX_DEVICES) return fail(DH_ERANGE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(DH_EHIP, std::string("hipGetSymbolAddress(scratch): ") +
                                     hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_dev[dev];
    return DH_OK;
}

}  // namespace

namespace {

// Steps first_step .. first_step+nsteps-1 (capped at max_total_steps) of a
// batch: nsteps_per_graph steps of sweep launches (launch(k, axis, stream))
// plus the step counter's advance, captured once per key into a cached
// hipGraph and replayed.
template <class Launch>
int batch_steps(const std::string &key, int D, int nsteps_per_graph, long long *step_base_dev,
                long long max_total_steps, long long first_step, long long nsteps, void *stream,
                Launch launch)
{
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    GraphRef entry = graph_lookup(key);
    if (!entry) {
        hipGraphExec_t exec = nullptr;
        hipStream_t cap;
        DH_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        DH_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        hipError_t el = hipSuccess;
        for (int k = 0; k < nsteps_per_graph && el == hipSuccess; ++k)
            for (int a = 0; a < D && el == hipSuccess; ++a) el = launch(k, a, cap);
        if (el == hipSuccess)
            hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, cap, step_base_dev,
                               (long long)nsteps_per_graph);
        hipGraph_t graph = nullptr;
        hipError_t ec = hipStreamEndCapture(cap, &graph);
        (void)hipStreamDestroy(cap);
        if (el != hipSuccess) ec = el;
        if (ec != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return fail(DH_EHIP, std::string("capture: ") + hipGetErrorString(ec));
        }
        hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess)
            return fail(DH_EHIP, std::string("instantiate: ") + hipGetErrorString(ei));
        entry = graph_insert(key, exec);
    }
    hipLaunchKernelGGL(k_set_ll, dim3(1), dim3(1), 0, st, step_base_dev, first_step);
    const long long last =
        first_step + nsteps < max_total_steps ? first_step + nsteps : max_total_steps;
    const long long reps = (last - first_step + nsteps_per_graph - 1) / nsteps_per_graph;
    for (long long r = 0; r < reps; ++r) DH_HIP(hipGraphLaunch(entry->exec, st));
    DH_HIP(entry->note_launch(st));
    return DH_OK;
}

}  // namespace

extern "C" {

int dh_abi_version(void) { return 6; }
size_t dh_sizeof_segment(void) { return sizeof(dh_segment); }
size_t dh_sizeof_instance(void) { return sizeof(dh_instance); }
size_t dh_lds_bytes(int D, int n, int waves, int chunk, int recompute, int gamma_global,
                    int delj)
{
    return nd_lds(D, n, waves, chunk, recompute, gamma_global != 0, delj != 0);
}

int dh_nd_groups(int D, int nmax, int waves, int chunk, int recompute, int gamma_global,
                 int groups, int ninst)
{
    if (int e = check_init()) return e;
    if (D < 1 || D > DH_MAXD || nmax < 3 || waves < 1 || waves > 8)
        return fail(DH_EINVAL, "bad launch description");
    return nd_groups(nd_lds(D, nmax, waves, chunk, recompute, gamma_global != 0), waves, groups,
                     ninst);
}
size_t dh_max_lds(void) { return g_max_lds; }
size_t dh_packed_lds_bytes(int D, int n, int waves, int slots)
{
    return packed_dyn_lds(D, n, waves, slots) + PACK_STATIC_LDS;
}

int dh_last_error(char *buf, int len)
{
    if (buf && len > 0) {
        std::strncpy(buf, g_err.c_str(), (size_t)len - 1);
        buf[len - 1] = 0;
    }
    return (int)g_err.size();
}

int dh_init(int device)
{
    DH_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    DH_HIP(hipGetDeviceProperties(&prop, device));
    g_num_cu = prop.multiProcessorCount;
    int maxsh = 0;
    if (hipDeviceGetAttribute(&maxsh, hipDeviceAttributeMaxSharedMemoryPerBlock, device) !=
            hipSuccess ||
        maxsh <= 0)
        maxsh = (int)prop.sharedMemPerBlock;
    g_max_lds = (size_t)maxsh;
    int wk = 0;
    if (hipDeviceGetAttribute(&wk, hipDeviceAttributeWallClockRate, device) != hipSuccess ||
        wk <= 0)
        wk = 100000;   // 100 MHz, the CDNA constant-rate counter
    g_wall_khz = wk;
    g_inited = true;
    double *sp = nullptr;
    return static_scratch(&sp);
}

int dh_integrate_nd(dh_instance *table_dev, int ninst, int D, int order, int nmax,
                    int *counter_dev, const dh_launch *cfg, void *stream)
{
    if (int e = check_init()) return e;
    if (D < 1 || D > DH_MAXD) return fail(DH_EINVAL, "D must be 1..5");
    if (order < 0 || order > 2) return fail(DH_EINVAL, "bad order");
    if (ninst <= 0) return DH_OK;
    if (!table_dev || !counter_dev || !cfg) return fail(DH_EINVAL, "null pointer");
    if (nmax < 3) return fail(DH_EINVAL, "n must be >= 3");
    const int waves = cfg->waves;
    if (waves < 1 || waves > 8) return fail(DH_EINVAL, "waves must be 1..8");
    const int chunk = cfg->chunk, recompute = cfg->recompute;
    const int gg = cfg->gamma_dev != nullptr, dj = cfg->delj != 0;
    if (chunk != 0 && chunk != 8) return fail(DH_EINVAL, "chunk must be 0 or 8");
    if (recompute && chunk != 8) return fail(DH_EINVAL, "recompute needs chunk 8");
    if (gg && (chunk || recompute))
        return fail(DH_EINVAL, "a global gamma workspace needs chunk 0 and recompute 0");
    if (dj && (chunk || recompute || gg))
        return fail(DH_EINVAL, "the delta trick needs chunk 0 without recompute or workspace");
    const void *fn = pick_nd(D, order, chunk, recompute, gg, dj);
    const size_t lds = nd_lds(D, nmax, waves, chunk, recompute, gg, dj);
    if (int e = set_lds(fn, lds)) return e;
    const int groups = nd_groups(lds, waves, cfg->groups, ninst);
    if (gg && cfg->gamma_doubles < (long long)groups * waves * 64 * nmax)
        return fail(DH_ERANGE, "gamma workspace too small: need " +
                                   std::to_string((long long)groups * waves * 64 * nmax) +
                                   " doubles");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, counter_dev, 0);
    long long ms = cfg->max_steps;
    double *gws = cfg->gamma_dev;
    long long ticks = budget_ticks(cfg);
    void *args[] = {&table_dev, &ninst, &counter_dev, &ms, &gws, &nmax, &ticks};
    DH_HIP(hipLaunchKernel(fn, dim3(groups), dim3(64 * waves), args, lds, st));
    return DH_OK;
}

int dh_integrate_packed(dh_instance *table_dev, int ninst, int D, int order, int n, int slots,
                        int *counter_dev, const dh_launch *cfg, void *stream)
{
    if (int e = check_init()) return e;
    if (D < 2 || D > DH_MAXD) return fail(DH_EINVAL, "packed launches need D = 2..5");
    if (order < 0 || order > 2) return fail(DH_EINVAL, "bad order");
    if (ninst <= 0) return DH_OK;
    if (!table_dev || !counter_dev || !cfg) return fail(DH_EINVAL, "null pointer");
    if (n < 3) return fail(DH_EINVAL, "n must be >= 3");
    if (slots < 1 || slots > DH_PACK_MAX)
        return fail(DH_EINVAL, "slots must be 1.." + std::to_string(DH_PACK_MAX));
    const int waves = cfg->waves;
    if (waves < 1 || waves > 8) return fail(DH_EINVAL, "waves must be 1..8");
    if (cfg->chunk != 8 || cfg->recompute || cfg->gamma_dev || cfg->delj)
        return fail(DH_EINVAL, "packed launches use chunk 8 without recompute, workspace or "
                               "delta trick");
    const void *fn = pick_packed(D, order);
    const size_t dyn = packed_dyn_lds(D, n, waves, slots);
    if (dyn + PACK_STATIC_LDS > g_max_lds)
        return fail(DH_ERANGE, "LDS request of " + std::to_string(dyn + PACK_STATIC_LDS) +
                                   " bytes exceeds " + std::to_string(g_max_lds));
    if (int e = set_lds(fn, dyn)) return e;
    const int groups =
        nd_groups(dyn + PACK_STATIC_LDS, waves, cfg->groups, (ninst + slots - 1) / slots);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, counter_dev, 0);
    long long ms = cfg->max_steps;
    void *args[] = {&table_dev, &ninst, &counter_dev, &ms, &slots, &n};
    DH_HIP(hipLaunchKernel(fn, dim3(groups), dim3(64 * waves), args, dyn, st));
    return DH_OK;
}

int dh_integrate_1d(dh_instance *table_dev, int ninst, int order, int nmax, int *counter_dev,
                    double *scratch_dev, long long scratch_doubles, const dh_launch *cfg,
                    void *stream)
{
    if (int e = check_init()) return e;
    if (order < 0 || order > 2) return fail(DH_EINVAL, "bad order");
    if (ninst <= 0) return DH_OK;
    if (!table_dev || !counter_dev || !cfg || !scratch_dev) return fail(DH_EINVAL, "null pointer");
    if (nmax < 3) return fail(DH_EINVAL, "n must be >= 3");
    const int waves = cfg->waves;
    if (waves < 1 || waves > 4) return fail(DH_EINVAL, "waves must be 1..4 for the 1D kernel");
    const int thr = 64 * waves;
    long long groups = cfg->groups;
    if (groups <= 0) groups = (ninst + thr - 1) / thr;
    const long long nlanes = groups * thr;
    const long long tables = cfg->delj ? 8 : 6;
    if (tables * nmax * nlanes > scratch_doubles)
        return fail(DH_ERANGE, "1D scratch too small: need " +
                                   std::to_string(tables * nmax * nlanes) + " doubles");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, counter_dev, 0);
    long long ms = cfg->max_steps;
    void *args[] = {&table_dev, &ninst, &counter_dev, &scratch_dev, (void *)&nlanes, &nmax, &ms};
    DH_HIP(hipLaunchKernel(pick_1d(order, cfg->delj), dim3((unsigned)groups), dim3(thr), args, 0,
                           st));
    return DH_OK;
}

long long dh_factored_slice_doubles(int D, int nmax)
{
    if (D < 2 || D > DH_MAXD || nmax < 3) return 0;
    return fz_slice(D, nmax);
}

}  // extern "C"

namespace {

// Checks shared by the two factored entry points (0 = go ahead, 1 = nothing
// to do, < 0 = error).
int factored_args(const dh_instance *table_dev, int ninst, int D, int order, int nmax,
                  const dh_launch *cfg)
{
    if (int e = check_init()) return e;
    if (D < 2 || D > DH_MAXD) return fail(DH_EINVAL, "D must be 2..5");
    if (order < 0 || order > 2) return fail(DH_EINVAL, "bad order");
    if (ninst <= 0) return 1;
    if (!table_dev || !cfg) return fail(DH_EINVAL, "null pointer");
    if (nmax < 3) return fail(DH_EINVAL, "n must be >= 3");
    if (cfg->delj) return fail(DH_EINVAL, "the delta trick is not available on this path");
    if (cfg->waves < 1 || cfg->waves > 8) return fail(DH_EINVAL, "waves must be 1..8");
    return DH_OK;
}

// Resets the queue head and launches the factored kernel over `groups`
// workgroups, each with its `slice` of the workspace `ws`.
int factored_launch(dh_instance *table_dev, int ninst, int D, int order, int nmax,
                    int *counter_dev, const dh_launch *cfg, double *ws, long long groups,
                    long long slice, void *stream)
{
    const void *fn = pick_fz(D, order);
    const size_t lds =
        sizeof(double) * (size_t)nmax * (LDS_GRID_TABLES + LDS_TABLES_PER_AXIS * (size_t)D);
    if (int e = set_lds(fn, lds)) return e;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, counter_dev, 0);
    long long ms = cfg->max_steps, ticks = budget_ticks(cfg);
    long long sl = slice;
    void *args[] = {&table_dev, &ninst, &counter_dev, &ms, &ws, &sl, &ticks};
    DH_HIP(hipLaunchKernel(fn, dim3((unsigned)groups), dim3(64 * cfg->waves), args, lds, st));
    return DH_OK;
}

}  // namespace

extern "C" {

int dh_integrate_factored(dh_instance *table_dev, int ninst, int D, int order, int nmax,
                          int *counter_dev, const dh_launch *cfg, void *stream)
{
    if (int e = factored_args(table_dev, ninst, D, order, nmax, cfg)) return e < 0 ? e : DH_OK;
    if (!counter_dev) return fail(DH_EINVAL, "null pointer");
    long long groups = cfg->groups > 0 ? cfg->groups : std::min<long long>(ninst, g_num_cu);
    const long long slice = fz_slice(D, nmax);
    if (!cfg->gamma_dev || cfg->gamma_doubles < groups * slice)
        return fail(DH_ERANGE, "factorisation workspace too small: need " +
                                   std::to_string(groups * slice) + " doubles");
    return factored_launch(table_dev, ninst, D, order, nmax, counter_dev, cfg, cfg->gamma_dev,
                           groups, slice, stream);
}

int dh_resident_rows(int n)
{
    size_t lds = 0;
    int waves = 0;
    return pick_rr(n, DH_ORDER_PRE, true, &lds, &waves) ? (n + RR_SPAN - 1) / RR_SPAN * RR_SPAN
                                                         : 0;
}

int dh_integrate_resident(dh_instance *table_dev, int ninst, int order, int n,
                          int *counter_dev, const dh_launch *cfg, void *stream)
{
    if (int e = check_init()) return e;
    if (order != DH_ORDER_TV && order != DH_ORDER_PRE)
        return fail(DH_EINVAL, "order must be TV or PRE (2 populations)");
    if (ninst <= 0) return DH_OK;
    if (!table_dev || !counter_dev || !cfg) return fail(DH_EINVAL, "null pointer");
    size_t lds = 0;
    int waves = 0;
    const void *fast = pick_rr(n, order, true, &lds, &waves);
    const void *exact = pick_rr(n, order, false, &lds, &waves);
    if (!fast || !exact)
        return fail(DH_ERANGE, "no register-resident kernel for n = " + std::to_string(n));
    if (int e = set_lds(fast, lds)) return e;
    if (int e = set_lds(exact, lds)) return e;
    // one workgroup per SIMD pair: each wave needs a SIMD's whole register file
    const int per_cu = std::max(1, 4 / waves);
    long long groups = cfg->groups > 0 ? cfg->groups
                                       : std::min<long long>(ninst, (long long)per_cu * g_num_cu);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    long long ms = cfg->max_steps, ticks = budget_ticks(cfg);
    void *args[] = {&table_dev, &ninst, &counter_dev, &ms, &ticks};
    // the fast kernel, then the exact one over the instances it handed over
    // (those whose quotients left the fast kernel's exact range)
    for (const void *fn : {fast, exact}) {
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, counter_dev, 0);
        DH_HIP(hipLaunchKernel(fn, dim3((unsigned)groups), dim3(64 * waves), args, lds, st));
    }
    return DH_OK;
}

long long dh_resident_latency_slice_doubles(int n)
{
    size_t lds = 0;
    int waves = 0;
    long long slice = 0;
    return pick_rf(n, DH_ORDER_PRE, true, &lds, &waves, &slice) ? slice : 0;
}

int dh_resident_latency_rows(int n)
{
    size_t lds = 0;
    int waves = 0, rows = 0;
    long long slice = 0;
    return pick_rf(n, DH_ORDER_PRE, true, &lds, &waves, &slice, &rows) ? rows : 0;
}

int dh_integrate_resident_latency(dh_instance *table_dev, int ninst, int order, int n,
                                  int *counter_dev, const dh_launch *cfg, void *stream)
{
    if (int e = check_init()) return e;
    if (order != DH_ORDER_TV && order != DH_ORDER_PRE)
        return fail(DH_EINVAL, "order must be TV or PRE (2 populations)");
    if (ninst <= 0) return DH_OK;
    if (!table_dev || !counter_dev || !cfg) return fail(DH_EINVAL, "null pointer");
    size_t lds = 0;
    int waves = 0;
    long long slice = 0;
    const void *fast = pick_rf(n, order, true, &lds, &waves, &slice);
    const void *exact = pick_rf(n, order, false, &lds, &waves, &slice);
    if (!fast || !exact)
        return fail(DH_ERANGE, "no register-resident kernel for n = " + std::to_string(n));
    if (!cfg->gamma_dev)
        return fail(DH_EINVAL, "gamma_dev must name the factorisation workspace "
                               "(dh_resident_latency_slice_doubles(n) doubles per workgroup)");
    if (int e = set_lds(fast, lds)) return e;
    if (int e = set_lds(exact, lds)) return e;
    // one wave per SIMD; the LDS tile may allow fewer workgroups per compute unit
    const int per_cu = std::max(1, std::min(4 / waves, (int)(g_max_lds / lds)));
    long long groups = cfg->groups > 0 ? cfg->groups
                                       : std::min<long long>(ninst, (long long)per_cu * g_num_cu);
    if (cfg->gamma_doubles < groups * slice)
        return fail(DH_ERANGE, "factorisation workspace too small: need " +
                                   std::to_string(groups * slice) + " doubles");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    long long ms = cfg->max_steps, ticks = budget_ticks(cfg);
    double *ws = cfg->gamma_dev;
    void *args[] = {&table_dev, &ninst, &counter_dev, &ms, &ticks, &ws, &slice};
    // the fast form, then the exact form over the instances it handed over
    // (or found marked by an earlier call)
    for (const void *fn : {fast, exact}) {
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, counter_dev, 0);
        DH_HIP(hipLaunchKernel(fn, dim3((unsigned)groups), dim3(64 * waves), args, lds, st));
    }
    return DH_OK;
}

long long dh_static_scratch_doubles(void) { return DH_STATIC_WORKSPACE; }

int dh_integrate_factored_static(dh_instance *table_dev, int ninst, int D, int order,
                                 int nmax, const dh_launch *cfg, void *stream)
{
    if (int e = factored_args(table_dev, ninst, D, order, nmax, cfg)) return e < 0 ? e : DH_OK;
    if (cfg->gamma_dev)
        return fail(DH_EINVAL, "gamma_dev must be NULL: this entry point uses the library's "
                               "static scratch");
    const long long slice = fz_slice(D, nmax);
    const long long fit = DH_STATIC_WORKSPACE / slice;
    if (fit < 1)
        return fail(DH_ERANGE, "one workgroup's factorisation workspace (" +
                                   std::to_string(slice) + " doubles for " + std::to_string(D) +
                                   " populations on " + std::to_string(nmax) +
                                   " points) exceeds the static scratch of " +
                                   std::to_string(DH_STATIC_WORKSPACE) + " doubles");
    long long groups;
    if (cfg->groups > 0) {
        groups = cfg->groups;
        if (groups > fit)
            return fail(DH_ERANGE, std::to_string(groups) + " workgroups need " +
                                       std::to_string(groups * slice) +
                                       " doubles of workspace; the static scratch holds " +
                                       std::to_string(DH_STATIC_WORKSPACE) + " (at most " +
                                       std::to_string(fit) + " workgroups)");
    } else {
        groups = std::min<long long>(std::min<long long>(ninst, g_num_cu), fit);
    }
    double *sp = nullptr;
    if (int e = static_scratch(&sp)) return e;
    int *counter = reinterpret_cast<int *>(sp);
    return factored_launch(table_dev, ninst, D, order, nmax, counter, cfg, sp + DH_STATIC_HEAD,
                           groups, slice, stream);
}

size_t dh_1d_wave_lds_bytes(int nmax)
{
    return (size_t)W1_ARRAYS * (size_t)(nmax + 2 * W1_PAD) * 8;
}

int dh_integrate_1d_wave(dh_instance *table_dev, int ninst, int order, int nmax,
                         int *counter_dev, const dh_launch *cfg, void *stream)
{
    if (int e = check_init()) return e;
    if (order < 0 || order > 2) return fail(DH_EINVAL, "bad order");
    if (ninst <= 0) return DH_OK;
    if (!table_dev || !counter_dev || !cfg) return fail(DH_EINVAL, "null pointer");
    if (nmax < 3) return fail(DH_EINVAL, "n must be >= 3");
    const void *fn = pick_1w(order, cfg->delj);
    const size_t lds = dh_1d_wave_lds_bytes(nmax);
    if (int e = set_lds(fn, lds)) return e;
    long long groups = cfg->groups;
    if (groups <= 0) groups = std::min<long long>(ninst, 32LL * g_num_cu);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, counter_dev, 0);
    long long ms = cfg->max_steps;
    long long ticks = budget_ticks(cfg);
    void *args[] = {&table_dev, &ninst, &counter_dev, &nmax, &ms, &ticks};
    DH_HIP(hipLaunchKernel(fn, dim3((unsigned)groups), dim3(64), args, lds, st));
    return DH_OK;
}

int dh_sweep_axis(const dh_instance *inst_host, int axis, int first_axis,
                  const long long *step_base_dev, long long step_offset,
                  const int *seg_of_step_dev, long long total_steps, int waves, void *stream)
{
    if (int e = check_init()) return e;
    if (!inst_host) return fail(DH_EINVAL, "null instance");
    const dh_instance I = *inst_host;
    if (I.D < 1 || I.D > DH_MAXD || I.order < 0 || I.order > 2 || axis < 0 || axis >= I.D)
        return fail(DH_EINVAL, "bad instance/axis");
    if (I.n < 3) return fail(DH_EINVAL, "n must be >= 3");
    if (waves < 1 || waves > 4) return fail(DH_EINVAL, "waves must be 1..4");
    long long nsys = 1;
    for (int q = 0; q < I.D - 1; ++q) nsys *= I.n;
    const long long tiles = (nsys + 63) / 64;
    long long groups = (tiles + waves - 1) / waves;
    const void *fn = pick_sweep(I.D, I.order);
    const size_t lds = sweep_lds(I.n, waves);
    if (int e = set_lds(fn, lds)) return e;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    dh_instance Ic = I;
    int a = axis, fa = first_axis;
    void *args[] = {&Ic, &a, &fa, (void *)&step_base_dev, &step_offset, (void *)&seg_of_step_dev,
                    &total_steps};
    DH_HIP(hipLaunchKernel(fn, dim3((unsigned)groups), dim3(64 * waves), args, lds, st));
    return DH_OK;
}

int dh_advance_steps(long long *step_base_dev, long long by, void *stream)
{
    if (int e = check_init()) return e;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, st, step_base_dev, by);
    DH_HIP(hipGetLastError());
    return DH_OK;
}

int dh_integrate_graph(const dh_instance *inst_host, long long *step_base_dev,
                       const int *seg_of_step_dev, long long total_steps, int nsteps_per_graph,
                       int waves, void *stream)
{
    return dh_integrate_graph_range(inst_host, step_base_dev, seg_of_step_dev, total_steps, 0,
                                    total_steps, nsteps_per_graph, waves, stream);
}

int dh_integrate_graph_range(const dh_instance *inst_host, long long *step_base_dev,
                             const int *seg_of_step_dev, long long total_steps,
                             long long first_step, long long nsteps, int nsteps_per_graph,
                             int waves, void *stream)
{
    if (int e = check_init()) return e;
    if (!inst_host || !step_base_dev || !seg_of_step_dev) return fail(DH_EINVAL, "null pointer");
    if (nsteps_per_graph < 1) return fail(DH_EINVAL, "nsteps_per_graph must be >= 1");
    if (first_step < 0 || nsteps < 0) return fail(DH_EINVAL, "negative step range");
    if (nsteps % nsteps_per_graph != 0 && first_step + nsteps < total_steps)
        return fail(DH_EINVAL, "nsteps must be a multiple of nsteps_per_graph or reach the end");
    if (total_steps <= 0 || nsteps == 0 || first_step >= total_steps) return DH_OK;
    const dh_instance I = *inst_host;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    // key: everything baked into the captured launches
    std::string key((const char *)&I, sizeof(I));
    key.append((const char *)&step_base_dev, sizeof(step_base_dev));
    key.append((const char *)&seg_of_step_dev, sizeof(seg_of_step_dev));
    key.append((const char *)&total_steps, sizeof(total_steps));
    key.append((const char *)&nsteps_per_graph, sizeof(int));
    key.append((const char *)&waves, sizeof(int));
    GraphRef entry = graph_lookup(key);
    if (!entry) {
        hipGraphExec_t exec = nullptr;
        hipStream_t cap;
        DH_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        DH_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        int rc = DH_OK;
        for (int k = 0; k < nsteps_per_graph && rc == DH_OK; ++k) {
            int first = 1;
            for (int a = 0; a < I.D && rc == DH_OK; ++a) {
                if ((I.frozen >> a) & 1) continue;
                rc = dh_sweep_axis(&I, a, first, step_base_dev, k, seg_of_step_dev, total_steps,
                                   waves, cap);
                first = 0;
            }
        }
        if (rc == DH_OK) hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, cap, step_base_dev,
                                            (long long)nsteps_per_graph);
        hipGraph_t graph = nullptr;
        hipError_t ec = hipStreamEndCapture(cap, &graph);
        (void)hipStreamDestroy(cap);
        if (rc != DH_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (ec != hipSuccess) return fail(DH_EHIP, std::string("capture: ") + hipGetErrorString(ec));
        hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess)
            return fail(DH_EHIP, std::string("instantiate: ") + hipGetErrorString(ei));
        entry = graph_insert(key, exec);
    }
    hipLaunchKernelGGL(k_set_ll, dim3(1), dim3(1), 0, st, step_base_dev, first_step);
    const long long last = first_step + nsteps < total_steps ? first_step + nsteps : total_steps;
    const long long reps = (last - first_step + nsteps_per_graph - 1) / nsteps_per_graph;
    for (long long r = 0; r < reps; ++r) DH_HIP(hipGraphLaunch(entry->exec, st));
    DH_HIP(entry->note_launch(st));
    return DH_OK;
}

size_t dh_sizeof_sweep_item(void) { return sizeof(dh_sweep_item); }

int dh_integrate_batch(const dh_sweep_item *items_dev, const int *wg_item_dev, int ninst,
                       int total_wgs, int D, int order, int nmax, int waves,
                       long long *step_base_dev, long long max_total_steps,
                       long long first_step, long long nsteps, int nsteps_per_graph,
                       void *stream)
{
    if (int e = check_init()) return e;
    if (!items_dev || !wg_item_dev || !step_base_dev) return fail(DH_EINVAL, "null pointer");
    if (D < 2 || D > DH_MAXD || order < 0 || order > 2) return fail(DH_EINVAL, "bad D/order");
    if (ninst <= 0 || total_wgs <= 0 || max_total_steps <= 0 || nsteps <= 0 ||
        first_step >= max_total_steps)
        return DH_OK;
    if (nmax < 3) return fail(DH_EINVAL, "n must be >= 3");
    if (waves < 1 || waves > 4) return fail(DH_EINVAL, "waves must be 1..4");
    if (nsteps_per_graph < 1) return fail(DH_EINVAL, "nsteps_per_graph must be >= 1");
    if (first_step < 0) return fail(DH_EINVAL, "negative first step");
    if (nsteps % nsteps_per_graph != 0 && first_step + nsteps < max_total_steps)
        return fail(DH_EINVAL, "nsteps must be a multiple of nsteps_per_graph or reach the end");
    // the full gamma column per lane (fewer instructions per row) while every
    // workgroup has a compute unit to itself and the column fits; with more
    // workgroups, checkpoints, whose smaller LDS lets several share a compute
    // unit (a batch of 288 3-population integrations, 4096 workgroups: 26%
    // more integrations/s)
    const size_t tables = (size_t)nmax * (LDS_GRID_TABLES + LDS_TABLES_PER_AXIS);
    const size_t lds0 = sizeof(double) * (tables + (size_t)waves * 64 * (size_t)nmax);
    const int ck = (lds0 <= g_max_lds && total_wgs <= g_num_cu) ? 0 : 8;
    const void *fn = pick_sweep_batch(D, order, ck);
    const size_t col = (size_t)((nmax - 1) / 8 + 1);
    const size_t lds = ck ? sizeof(double) * (tables + (size_t)waves * 64 * (col + STAGE_C + 1))
                          : lds0;
    if (int e = set_lds(fn, lds)) return e;
    std::string key("batch");
    key.append((const char *)&items_dev, sizeof(items_dev));
    key.append((const char *)&wg_item_dev, sizeof(wg_item_dev));
    key.append((const char *)&step_base_dev, sizeof(step_base_dev));
    const int ints[] = {ninst, total_wgs, D, order, nmax, waves, nsteps_per_graph};
    key.append((const char *)ints, sizeof(ints));
    return batch_steps(key, D, nsteps_per_graph, step_base_dev, max_total_steps, first_step,
                       nsteps, stream, [&](int k, int a, hipStream_t cap) {
                           long long off = k;
                           int ax = a;
                           void *args[] = {(void *)&items_dev, (void *)&wg_item_dev, &ax,
                                           (void *)&step_base_dev, &off};
                           return hipLaunchKernel(fn, dim3((unsigned)total_wgs),
                                                  dim3(64 * waves), args, lds, cap);
                       });
}

int dh_sample_axis(const double *in_dev, long long P, int n, const double *xx_dev, int m,
                   const double *db1_dev, const double *db2_dev, double *work_dev,
                   double *out_dev, void *stream)
{
    if (int e = check_init()) return e;
    if (!in_dev || !xx_dev || !db1_dev || !db2_dev || !work_dev || !out_dev)
        return fail(DH_EINVAL, "null pointer");
    if (P <= 0 || n < 2 || m < 0) return fail(DH_EINVAL, "bad sampling shape");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const long long tot = P * (n - 1);
    double *sT = work_dev, *cT = work_dev + tot;
    const int B = 256;
    hipLaunchKernelGGL(k_sample_prep, dim3((unsigned)((tot + B - 1) / B)), dim3(B), 0, st, in_dev,
                       P, n, xx_dev, m, sT, cT);
    const long long outn = (long long)(m + 1) * P;
    hipLaunchKernelGGL(k_sample_contract, dim3((unsigned)((outn + B - 1) / B)), dim3(B), 0, st,
                       sT, cT, P, n, m, db1_dev, db2_dev, out_dev);
    DH_HIP(hipGetLastError());
    return DH_OK;
}

int dh_contract_axis(const double *in_dev, long long P, int n, const double *A_dev, int m1,
                     double *out_dev, void *stream)
{
    if (int e = check_init()) return e;
    if (!in_dev || !A_dev || !out_dev) return fail(DH_EINVAL, "null pointer");
    if (P <= 0 || n < 1 || m1 < 1) return fail(DH_EINVAL, "bad contraction shape");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const long long outn = (long long)m1 * P;
    const int B = 256;
    hipLaunchKernelGGL(k_contract_axis, dim3((unsigned)((outn + B - 1) / B)), dim3(B), 0, st,
                       in_dev, P, n, A_dev, m1, out_dev);
    DH_HIP(hipGetLastError());
    return DH_OK;
}

}  // extern "C"
