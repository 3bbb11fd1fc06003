// Dense complex-symmetric factorisation Z = L D L^T of the free-loop impedance on the device, and
// the solves with it: the direct (factor-and-refine) port solve of pq_loops_solve_direct.
#pragma once
#include <hip/hip_runtime.h>

namespace pq {

constexpr int LDL_BLOCK = 64;   // panel width, diagonal block, trailing-update tile

// The factor's storage: planes Fr, Fi (n x ldf FP32, row-major, ldf a multiple of 64; the lower
// triangle holds L below the diagonal and D on it), the inverses of the unit lower diagonal blocks
// Lir, Lii (ceil(n / 64) blocks of 64 x 64), the trailing update's right factor Wr, Wi (two
// buffers of n x 128, alternating between 128-wide super-panels). The factorisation also uses a second stream and
// two events of the caller's (the lookahead: the bulk of one panel's trailing update runs beside
// the next panel's diagonal block and panel).
struct LdlFactor {
    float *Fr, *Fi;
    long long ldf;
    int n;
    float *Lir, *Lii;
    float *Wr[2], *Wi[2];
    hipStream_t aux;
    hipEvent_t ready, done;
    // `batch` matrices of the same order factored together (the kernels' blockIdx.z), matrix z's
    // planes at Fr/Fi + z sF, its block inverses at Lir/Lii + z sX, its W buffers at Wr/Wi + z sW
    int batch = 1;
    long long sF = 0, sX = 0, sW = 0;
    // band > 0: every entry lies within `band` of the diagonal (|i - j| <= band), so the factor does
    // too, and the factorisation and the sweeps skip the tiles beyond it (storage stays dense)
    int band = 0;
};

// The factorisation's second stream (the bulk trailing updates): a plain non-blocking stream. With
// PCBQS_AMD_LDL_CUMASK=1 it runs on all CUs but eight, one taken from every 32 (CU 32 i + i, one per
// chiplet under either CU numbering), leaving them to the diagonal blocks and panels on the critical
// path; that gained 1-3 % in a process with few streams, but a masked stream needs a hardware queue
// of its own and cost 45 % in one with several (the queues are then shared), so it is off by default.
void ldl_create_aux(hipStream_t *aux);

// F = Rl + j omega Ll on the first n loops (lower triangle): Ll FP32 (row stride ldl), Rl CSR FP64.
void ldl_form(hipStream_t st, const LdlFactor &f, const float *Ll, long long ldl, const int *rptr, const int *rcol,
              const double *rval, double omega);

// Factor in place on stream st (the work is complete in st's order when it returns; f.aux has no
// pending work left behind; with f.aux null everything runs on st, in order, as graph capture needs); bad[z] (f.batch device ints, cleared by the caller) is set when a pivot
// of matrix z is zero or not finite.
void ldl_factor(hipStream_t st, const LdlFactor &f, int *bad);

// X = Z^-1 B for P complex columns (n, 2P) FP64 (real parts first), FP32 factor, FP64 arithmetic.
// Y, Ys: (n, 2P) FP64 work arrays; B, Y, Ys and X do not overlap. P is 1..8 or 16. With
// f.batch > 1, B, Y, Ys and X hold the batch's systems one after another (n x 2P each) and all are
// solved in the same launches.
void ldl_solve(hipStream_t st, const LdlFactor &f, int P, const double *B, double *Y, double *Ys, double *X);

}  // namespace pq
