// partition_solver.hip -- workgroup-parallel tridiagonal line solver for gfx950.
//
// Solves many independent tridiagonal systems (periodic or not) of n unknowns each,
//     b_k x_{k-1} + a_k x_k + c_k x_{k+1} = f_k,   k = 0..n-1,
// with x_{-1} = x_{n-1} and x_n = x_0 when periodic (corners b_0 and c_{n-1}, the layout
// of Math::TridiagonalSolverPeriodic, src/Math.cpp:68-118) and b_0 = c_{n-1} = 0 otherwise
// (Math::TridiagonalSolver, src/Math.cpp:15-65).
//
// Each line is split into S segments (S a power of two), one per thread:
//   1. Two-sided elimination inside the segment, in registers: the downward pass removes
//      the sub-diagonal and leaves a fill-in column on the segment's first unknown, the
//      upward pass removes the super-diagonal and leaves one on its last unknown.
//   2. The first/last unknowns of all segments form a cyclic tridiagonal system of
//      N = 2S rows, solved by parallel cyclic reduction in LDS: log2(N)-1 reduction steps
//      leave rows i and i+N/2 coupled only to each other, then a 2x2 solve.
//   3. Interior unknowns follow from the two boundary values.
// The matrices are diagonally dominant (the 1/dt term), so no pivoting is needed. The
// rounding differs from the serial recurrences of the reference; results agree to a few
// ulps per solve.
#include "hydro_device.h"
#include "hydro_kernels.h"

namespace m70 {


struct LineGeom {
    long ld;        // COLS: row stride of the arrays; ROWS: line stride
    long ldy;       // output stride (as ld, for Y)
    int n;          // unknowns per line
    int nL;         // number of lines
    int S;          // segments per line (power of two, n >= 2S, n <= SEG_M*S)
    int log2S;
    int lines_per_block;
    int periodic;
    int xcd_remap;  // remap block index so neighbouring tiles share an XCD's L2
};

template <bool COLS>
__device__ __forceinline__ long elem_index(const LineGeom &g, int line, int k, long ld)
{
    return COLS ? (long)k * ld + line : (long)line * ld + k;
}

// One block: lines_per_block lines x S segments. COLS: lines are columns of row-major
// arrays (element (k, line) at k*ld + line) and thread t maps to (line t % L, segment
// t / L) so that a wavefront reads runs of consecutive columns; ROWS: lines are rows
// (element (line, k) at line*ld + k) and t maps to (line t / S, segment t % S).
template <bool COLS, int BLOCK, int SEG_M>
__global__ void __launch_bounds__(BLOCK) k_partition_solve(const double *__restrict__ A,
                                                           const double *__restrict__ B,
                                                           const double *__restrict__ C,
                                                           const double *__restrict__ F, double *__restrict__ Y,
                                                           LineGeom g)
{
    __shared__ double red[4][BLOCK * 2];   // reduced rows: lower, diag, upper, rhs
    const int t = threadIdx.x;
    const int S = g.S, L = g.lines_per_block;
    int blk = blockIdx.x;
    if (g.xcd_remap) {
        const int nb = gridDim.x;
        if ((nb & 7) == 0) blk = (blk & 7) * (nb >> 3) + (blk >> 3);
    }
    const int lloc = COLS ? (t % L) : (t / S);
    const int seg = COLS ? (t / L) : (t % S);
    const int line = blk * L + lloc;
    const bool active = (t < L * S) && line < g.nL;
    const int n = g.n;
    const int k0 = (int)(((long)seg * n) / S);
    const int k1 = (int)(((long)(seg + 1) * n) / S);
    const int m = k1 - k0;   // 2..SEG_M

    double lo[SEG_M], di[SEG_M], up[SEG_M], rh[SEG_M];
    if (active) {
#pragma unroll
        for (int q = 0; q < SEG_M; ++q) {
            if (q < m) {
                const long e = elem_index<COLS>(g, line, k0 + q, g.ld);
                di[q] = A[e];
                lo[q] = B[e];
                up[q] = C[e];
                rh[q] = F[e];
            }
        }
        if (!g.periodic) {
            if (k0 == 0) lo[0] = 0.0;
            if (k1 == n) up[m - 1] = 0.0;
        }
    }
    // 1a. downward pass: row q -> s_q x_0 + d_q x_q + up_q x_{q+1} = g_q (q >= 1).
    //     Registers are recycled as the pass consumes them: lo[q] becomes s_q, di[q]
    //     becomes 1/d_q, rh[q] becomes g_q.
    double A_F = 0, C_F = 0, R_F = 0, A_L = 0, B_L = 0, R_L = 0, lo0 = 0, upL = 0;
    if (active) {
        lo0 = lo[0];
        double dLast = di[1];         // d_{m-1}, kept un-inverted for the reduced row
        di[1] = 1.0 / di[1];          // s_1 = lo[1] already in place
#pragma unroll
        for (int q = 2; q < SEG_M; ++q) {
            if (q < m) {
                const double tq = lo[q] * di[q - 1];
                lo[q] = -tq * lo[q - 1];
                const double dq = di[q] - tq * up[q - 1];
                rh[q] = rh[q] - tq * rh[q - 1];
                di[q] = 1.0 / dq;
                if (q == m - 1) dLast = dq;
            }
        }
        // row L (q = m-1): A_L x_0 + B_L x_{m-1} + upL x_m = R_L
        {
            double sL = lo[1], gL = rh[1], uL = up[1];
#pragma unroll
            for (int q = 2; q < SEG_M; ++q)
                if (q == m - 1) { sL = lo[q]; gL = rh[q]; uL = up[q]; }
            A_L = sL; B_L = dLast; R_L = gL; upL = uL;
        }
        // 1b. upward pass: row q -> s'_q x_0 + d_q x_q + q_q x_{m-1} = g'_q for
        //     q = m-2..1; up[q] becomes the fill-in q_q on x_{m-1}.
#pragma unroll
        for (int q = SEG_M - 2; q >= 1; --q) {
            if (q < m - 2) {
                const double tq = up[q] * di[q + 1];
                lo[q] = lo[q] - tq * lo[q + 1];
                up[q] = -tq * up[q + 1];
                rh[q] = rh[q] - tq * rh[q + 1];
            }
        }
        // row F (q = 0): lo0 x_{-1} + A_F x_0 + C_F x_{m-1} = R_F
        if (m >= 3) {
            const double tq = up[0] * di[1];
            A_F = di[0] - tq * lo[1];
            C_F = -tq * up[1];
            R_F = rh[0] - tq * rh[1];
        } else {
            A_F = di[0];
            C_F = up[0];
            R_F = rh[0];
        }
    }

    // 2. cyclic PCR on the 2S reduced rows of each line, in LDS.
    const int N = 2 * S;
    const int base = lloc * N;   // rows of this line in red[][]
    const int rF = base + 2 * seg, rL = rF + 1;
    if (active) {
        red[0][rF] = lo0; red[1][rF] = A_F; red[2][rF] = C_F; red[3][rF] = R_F;
        red[0][rL] = A_L; red[1][rL] = B_L; red[2][rL] = upL; red[3][rL] = R_L;
    }
    __syncthreads();
    for (int st = 1; st < S; st <<= 1) {   // strides 1, 2, ..., N/4
        double na[2], nb[2], nc[2], nd[2];
        if (active) {
#pragma unroll
            for (int w = 0; w < 2; ++w) {
                const int i = 2 * seg + w;
                const int im = base + ((i - st) & (N - 1)), ip = base + ((i + st) & (N - 1));
                const int ii = base + i;
                const double k1m = red[0][ii] / red[1][im];
                const double k2p = red[2][ii] / red[1][ip];
                na[w] = -red[0][im] * k1m;
                nc[w] = -red[2][ip] * k2p;
                nb[w] = red[1][ii] - red[2][im] * k1m - red[0][ip] * k2p;
                nd[w] = red[3][ii] - red[3][im] * k1m - red[3][ip] * k2p;
            }
        }
        __syncthreads();
        if (active) {
#pragma unroll
            for (int w = 0; w < 2; ++w) {
                const int ii = base + 2 * seg + w;
                red[0][ii] = na[w]; red[1][ii] = nb[w]; red[2][ii] = nc[w]; red[3][ii] = nd[w];
            }
        }
        __syncthreads();
    }
    double xF = 0, xL = 0;
    if (active) {
        // rows i and j = i + N/2 are coupled only to each other:
        //   b_i x_i + e_i x_j = d_i,  e_j x_i + b_j x_j = d_j,  e = lower + upper
        double xs[2];
#pragma unroll
        for (int w = 0; w < 2; ++w) {
            const int i = 2 * seg + w;
            const int ii = base + i, jj = base + ((i + S) & (N - 1));
            const double bi = red[1][ii], ei = red[0][ii] + red[2][ii], di_ = red[3][ii];
            const double bj = red[1][jj], ej = red[0][jj] + red[2][jj], dj = red[3][jj];
            xs[w] = (di_ * bj - ei * dj) / (bi * bj - ei * ej);
        }
        xF = xs[0];
        xL = xs[1];
    }

    // 3. back-substitution and store
    if (active) {
#pragma unroll
        for (int q = 0; q < SEG_M; ++q) {
            if (q < m) {
                double x;
                if (q == 0) x = xF;
                else if (q == m - 1) x = xL;
                else x = (rh[q] - lo[q] * xF - up[q] * xL) * di[q];
                Y[elem_index<COLS>(g, line, k0 + q, g.ldy)] = x;
            }
        }
    }
}

// Picks S (power of two, n >= 2S, ceil(n/S) <= M) for a line length n; 0 if none.
static int pick_segments(int n, int M, int smax)
{
    int S = 1;
    while (S < smax && (n + S - 1) / S > M) S <<= 1;
    while (S > 1 && n < 2 * S) S >>= 1;
    if ((n + S - 1) / S > M || n < 2) return 0;
    return S;
}

bool partition_supported(int n) { return pick_segments(n, 16, 256) > 0 || pick_segments(n, 8, 512) > 0; }

template <bool COLS, int BLOCK, int M>
static hipError_t launch_partition(hipStream_t st, const double *A, const double *B, const double *C,
                                   const double *F, double *Y, long ld, long ldy, int n, int nL, bool periodic)
{
    LineGeom g{};
    g.ld = ld; g.ldy = ldy; g.n = n; g.nL = nL; g.periodic = periodic ? 1 : 0;
    // columns: keep at least 4 lines per block so a row access covers 32 contiguous bytes
    g.S = pick_segments(n, M, COLS ? BLOCK / 4 : BLOCK);
    if (g.S == 0) return hipErrorInvalidValue;
    int lg = 0;
    while ((1 << lg) < g.S) ++lg;
    g.log2S = lg;
    g.lines_per_block = BLOCK / g.S;
    const int blocks = (nL + g.lines_per_block - 1) / g.lines_per_block;
    g.xcd_remap = COLS ? 1 : 0;
    hipLaunchKernelGGL((k_partition_solve<COLS, BLOCK, M>), dim3((unsigned)blocks), dim3(BLOCK), 0, st, A, B, C, F,
                       Y, g);
    return hipGetLastError();
}

// tile: unknowns per thread (8 or 16); 0 picks the default for the layout.
hipError_t launch_partition_cols(hipStream_t st, const double *A, const double *B, const double *C, const double *F,
                                 double *Y, int n, int m, long ld, int tile)
{
    if (tile == 16 && pick_segments(n, 16, 64) > 0)
        return launch_partition<true, 256, 16>(st, A, B, C, F, Y, ld, ld, n, m, false);
    if (pick_segments(n, 8, 256) > 0) return launch_partition<true, 1024, 8>(st, A, B, C, F, Y, ld, ld, n, m, false);
    return launch_partition<true, 256, 16>(st, A, B, C, F, Y, ld, ld, n, m, false);
}

hipError_t launch_partition_rows(hipStream_t st, const double *A, const double *B, const double *C, const double *F,
                                 long ld, double *Y, long ldy, int nL, int n, bool periodic, int tile)
{
    if (tile == 8 && pick_segments(n, 8, 512) > 0)
        return launch_partition<false, 512, 8>(st, A, B, C, F, Y, ld, ldy, n, nL, periodic);
    if (pick_segments(n, 16, 256) > 0)
        return launch_partition<false, 256, 16>(st, A, B, C, F, Y, ld, ldy, n, nL, periodic);
    return launch_partition<false, 512, 8>(st, A, B, C, F, Y, ld, ldy, n, nL, periodic);
}

}  // namespace m70
