/*
 * pcbqs_amd.h -- C ABI of libpcbqs_hip.so, the MI355X (gfx950) static-fill and
 * loop-solve library behind the `mi355x` pcbqs.static_backends plugin.
 *
 * Conventions
 *   - Every function returns 0 on success and a nonzero pq_status otherwise;
 *     nothing throws across the ABI. pq_last_error() gives the message of the
 *     calling thread's last failure.
 *   - Host pointers are plain C arrays, row-major. Device pointers (named *_dev)
 *     are addresses returned by pq_malloc (or any HIP allocation on the same
 *     device). `stream` is a hipStream_t (NULL: the library's own stream).
 *   - A handle is not thread-safe; calls on one handle must not overlap.
 *
 * The reference's boundary for this path is a Python protocol (a factory
 * returning an object with evaluate / fill_symmetric / parallel_tiles); each
 * entry point below cites the reference interface it replaces.
 */
#ifndef PCBQS_AMD_H
#define PCBQS_AMD_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum pq_status {
    PQ_OK = 0,
    PQ_EINVAL = 1,      /* invalid argument (ValueError in Python)                  */
    PQ_ENONFINITE = 2,  /* non-finite moments (FloatingPointError)                  */
    PQ_EDEVICE = 3,     /* HIP runtime / device failure (RuntimeError)              */
    PQ_ENOMEM = 4,      /* device allocation failed (MemoryError)                   */
    PQ_ECONVERGE = 5    /* iterative solve did not converge (ConvergenceError)      */
};

/* Read-only triangle geometry (float64, C order), the arrays of a prepared
 * FillGeometry: src/pcbqs/static_backends.py:105-130 (prepare), the level
 * tuples of src/pcbqs/_numerics/sing.py:56-69, and the dict the reference core
 * reads in src/pcbqs/core/static_kernels.hpp:16-43 (static_geometry). */
typedef struct pq_geometry {
    int32_t nt;            /* triangles                                         */
    const double *V;       /* (nt,3,2) vertices, centred                        */
    const double *L;       /* (nt,3)   edge lengths                             */
    const double *tau;     /* (nt,3,2) unit edge tangents                       */
    const double *un;      /* (nt,3,2) outward edge normals                     */
    const double *area;    /* (nt)                                              */
    const double *P3;      /* (nt,3,2) far-rule points                          */
    const double *qp[3];   /* (nt,nq[k],2) observation points of level k        */
    const double *qw[3];   /* (nq[k]) observation weights of level k            */
    int32_t nq[3];         /* 6 * 4^k (a level not prepared repeats level 0)    */
    const double *z;       /* (nt) sheet height of each triangle                */
    const double *cent;    /* (nt,2) centroids                                  */
    const double *radius;  /* (nt) bounding radii                               */
} pq_geometry;

typedef struct pq_fill pq_fill;   /* device copy of one geometry + scratch      */

/* ---- runtime ------------------------------------------------------------ */
const char *pq_last_error(void);
int pq_device_count(int *count);
int pq_malloc(int device, size_t bytes, void **ptr_dev);
int pq_free(void *ptr_dev);
int pq_memcpy_h2d(void *dst_dev, const void *src, size_t bytes);
int pq_memcpy_d2h(void *dst, const void *src_dev, size_t bytes);
int pq_memcpy_d2d(void *dst_dev, const void *src_dev, size_t bytes);
/* count elements of elem_bytes each, stride_bytes apart on the device, packed
 * on the host (e.g. a matrix diagonal: stride = (ld + 1) * element size). */
int pq_memcpy_d2h_strided(void *dst, const void *src_dev, size_t elem_bytes, size_t stride_bytes, size_t count);
int pq_memset(void *dst_dev, int value, size_t bytes);
int pq_synchronize(int device);
/* A non-blocking stream of its own for a caller that runs several pipelines
 * at once (one per board), its synchronisation and release. */
int pq_stream_create(int device, void **stream);
int pq_stream_destroy(void *stream);
int pq_stream_synchronize(void *stream);
/* The store ceiling of the fill's output pattern: nplanes (1..4) FP32 planes of
 * n x n at leading dimension ld in planes_dev, each 64 x 64 tile written as
 * 256-byte row segments with non-temporal stores (value r - c), as the FP32
 * fill stores them, without the pair arithmetic; *ms is the mean of three
 * timed launches. Measurement only: no counterpart in the reference. */
int pq_store_probe(void *planes_dev, int64_t n, int64_t ld, int nplanes, double *ms);

/* ---- fill --------------------------------------------------------------- */

/* Upload one geometry to `device` once (geometry stays read-only).
 * Replaces CupyBackend.__init__ (src/pcbqs/static_cupy.py:73-107) / the
 * factory make_backend(geometry, workers, chunk, device)
 * (docs/static-backends.md:157-193). */
int pq_fill_create(const pq_geometry *g, int device, pq_fill **out);
int pq_fill_destroy(pq_fill *f);

/* Whole reciprocal FP64 fill into caller-owned host (nt,nt) matrices
 * planes[0..3] = M0, Mrx, Mry, Mrr (planes[0] only when scalar_only), with
 * near = INFINITY for the all-analytic rule; writes the unordered analytic
 * pair count. Replaces CupyBackend.fill_symmetric
 * (src/pcbqs/static_cupy.py:127-197), called from static_moments
 * (src/pcbqs/static.py:35-42); same result as static_fill
 * (src/pcbqs/core/static_kernels.hpp:184-252). PQ_ENONFINITE when any moment
 * is not finite. */
int pq_fill_symmetric(pq_fill *f, double near, int outer_levels, int scalar_only,
                      double *const planes[4], int64_t *analytic);

/* Explicit pair list: p, q int32, d float64, kind int32 in {-1, 0, 1, 2}
 * (levels must be prepared); out float64 (n,6) [M0, Mro_x, Mro_y, Mrs_x,
 * Mrs_y, Mrr] or (n,1) when scalar_only, host memory, synchronized.
 * Replaces CupyBackend.evaluate (src/pcbqs/static_cupy.py:108-125), the
 * backend side of FillEngine.evaluate (src/pcbqs/static_backends.py:300-336);
 * same result as static_pairs (src/pcbqs/core/static_kernels.hpp:147-172). */
int pq_pairs(pq_fill *f, const int32_t *p, const int32_t *q, const double *d,
             const int32_t *kind, int64_t n, int scalar_only, double *out);

/* pq_pairs without any device allocation: the same arguments and the same
 * result, bit for bit, with its staging buffers (the pair arrays, the kind
 * order and the moments) in a fixed device array of the library (about 75 MB
 * per device, held until the process ends) instead of memory allocated per
 * call. At most 1048576 pairs per call (PQ_EINVAL beyond). Calls on one device
 * share that array: calls that overlap on one device race, so callers
 * serialise them. No counterpart in the reference: for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). */
int pq_pairs_noalloc(pq_fill *f, const int32_t *p, const int32_t *q, const double *d,
                     const int32_t *kind, int64_t n, int scalar_only, double *out);

/* Device-resident fill: planes_dev holds 1 or 4 planes of nt*nt elements,
 * plane k at planes_dev + k*nt*nt, row-major. precision 64: reference layout
 * (moments about the centred origin, FP64 throughout). precision 32: FP32
 * planes, far pairs in FP32, near pairs in FP64, first moments taken about each
 * row/column triangle's own centroid ("local frame": Mrx[r,c] = sum g (x_r -
 * c_r), Mrr[r,c] = sum g (x_r - c_r).(x_c - c_c)) so that the RWG combination
 * does not cancel. Same rule and pair set as static_fill
 * (src/pcbqs/core/static_kernels.hpp:184-252); asynchronous on `stream`
 * except that it returns the analytic count (one small copy). */
int pq_fill_device(pq_fill *f, double near, int outer_levels, int scalar_only, int precision,
                   void *planes_dev, int64_t ld, int64_t *analytic, void *stream);

/* Number of pair-kernel launches and timings of the last pq_fill_device call:
 * far kernel milliseconds, near kernel milliseconds (HIP events on the launch
 * stream), near pairs per level. Not in the reference (measurement only). */
int pq_fill_last_timing(pq_fill *f, double *far_ms, double *near_ms, int64_t near_pairs[2]);

/* ---- RWG combination ---------------------------------------------------- */

/* L[m,n] = factor s_m s_n sum_{i,j} a_mi a_nj [Mrr(p,q) - v_nj.Mr(p,q)
 *          - v_mi.Mr(q,p) + v_mi.v_nj M0(p,q)], p = tri[i][m], q = tri[j][n],
 * written symmetrically into L_dev (ne rows, leading dimension ldl >= ne; 0:
 * ne). planes_dev (leading dimension ld, 0: nt) as produced by
 * pq_fill_device with the same precision (32: local frame; cent_dev (nt,2)
 * float64 centroids are then required). tri_dev (2,ne) int32, alpha_dev (2,ne),
 * vertex_dev (2,ne,2), scale_dev (ne) float64. out_precision 32 or 64.
 * Replaces static_combine (src/pcbqs/core/static_kernels.hpp:257-310) as
 * called by edge_inductance (src/pcbqs/static.py:103-128). */
int pq_combine_device(int precision, const void *planes_dev, int64_t nt, int64_t ld, const double *cent_dev,
                      const int32_t *tri_dev, const double *alpha_dev, const double *vertex_dev,
                      const double *scale_dev, int64_t ne, double factor, int out_precision,
                      void *L_dev, int64_t ldl, void *stream);

/* ---- loop-space port solve --------------------------------------------- */

/* Potential coefficients of the charge nodes from the device M0 plane (FP32 or
 * FP64, (nt, ld)): P[a][b] = factor / (A_a A_b) sum_{p in a, q in b} M0[p][q]
 * into P_dev (n, ldp) FP32 or FP64, node a's triangles node_tris[node_ptr[a]
 * .. node_ptr[a+1]) (device arrays), A = node_area (device). factor =
 * 1 / (4 pi eps0 er). Replaces MQS._set_potential (src/pcbqs/mqs.py:532-556),
 * P = W^T M0 W / (4 pi eps0 er), without the side-charge panels. */
int pq_potential_device(int precision, const void *M0_dev, int64_t nt, int64_t ld, int64_t n,
                        const int32_t *node_ptr, const int32_t *node_tris, const double *node_area, double factor,
                        int out_precision, void *P_dev, int64_t ldp, void *stream);

/* X = A^-1 B for a real symmetric positive definite A (n x n FP64 on the device,
 * row stride lda; its lower triangle is read) and nrhs right-hand sides B
 * (n x nrhs FP64 host, row-major; X likewise): A factored as L D L^T in FP32 on
 * the device, the solution refined in FP64 against A until every column's
 * relative residual is at most tol, at most max_steps corrections, stopping
 * once a correction no longer halves it (above 1e-10: PQ_ECONVERGE). *steps:
 * the most corrections any column group took; *residual: the largest final
 * relative residual; *factor_ms: device time of forming and factoring. A
 * non-positive-definite A (zero pivot) returns PQ_EINVAL. Replaces
 * solve_symmetric's role in capacitance_matrix (src/pcbqs/peec.py:95-123):
 * C = B^T P^-1 B on the potential matrix P. */
int pq_spd_solve(int device, int n, const double *A_dev, int64_t lda, int nrhs, const double *B_host, double tol,
                 int max_steps, double *X_host, int64_t *steps, double *residual, double *factor_ms);

typedef struct pq_csym pq_csym;     /* a factored complex-symmetric matrix        */

/* Factor the complex-symmetric n x n matrix given by its lower-triangle
 * entries (rows >= cols, unique; vals interleaved complex FP64) on the device:
 * FP32 L D L^T without pivoting (the real part positive definite: a graph
 * Laplacian A D^-1 A^T with Re 1/d > 0), keeping the entries in FP64 for
 * refinement. When every entry lies within a band (row - col <= b with
 * b + 128 < n: the Laplacian in a bandwidth-reducing order), the factorisation
 * and the triangular sweeps work inside the band only (O(n b^2) work; the
 * storage stays n x n). *factor_ms: device time of forming and factoring;
 * *bytes: the factor's storage. A zero pivot returns PQ_EINVAL; a factor
 * larger than the device's free memory returns PQ_ENOMEM with the bytes it
 * needs (the dense storage limits n to about 140k on 288 GB). Replaces the
 * per-frequency splu of the graph preconditioner's Laplacian
 * (src/pcbqs/mqs_precond.py:50-51). */
int pq_csym_create(int device, int n, int64_t nnz, const int32_t *rows, const int32_t *cols, const double *vals,
                   pq_csym **out, double *factor_ms, int64_t *bytes);
/* X = A^-1 B, B and X (n x nrhs) interleaved complex FP64 host arrays: FP32
 * factor solves, FP64 residuals against the entries, corrections until the
 * relative residual of every column is <= tol (at most max_steps, stopping
 * when it no longer halves); a final residual above max(tol, 1e-10) returns
 * PQ_ECONVERGE. Replaces lu.solve of the same preconditioner
 * (src/pcbqs/mqs_precond.py:57). */
int pq_csym_solve(pq_csym *h, int nrhs, const double *B, double tol, int max_steps, double *X, int64_t *steps,
                  double *residual);
int pq_csym_destroy(pq_csym *h);

typedef struct pq_loops pq_loops;   /* a board's loop operator on one device      */

/* Project the branch inductance onto the loop basis on the device and keep the
 * loop operator: Ll = C^T Lb C (FP64 accumulation and storage, symmetrised,
 * plus a copy rounded to FP32 for the Krylov products, the preconditioners
 * and the FP32 factor) and Rl (CSR, FP64). C is given in CSC form over loops
 * (c_colptr (n_loops+1), c_rows, c_sign, c_nnz entries; branch rows >=
 * n_branch_rows carry no inductance); Lb_dev is the (n_branch_rows, ldb)
 * branch inductance (RWG edges then via segments), FP32 or FP64 as
 * lb_precision (32 or 64) says. Loops are ordered free loops first (n_free),
 * then the port loops. Replaces the dense loop matrices of MQS._assemble /
 * _project_inductance (src/pcbqs/mqs.py:639-665). */
int pq_loops_create(int device, int n_loops, int n_free, const int32_t *c_colptr, const int32_t *c_rows,
                    const float *c_sign, int64_t c_nnz, const void *Lb_dev, int lb_precision, int64_t ldb,
                    int n_branch_rows, const int32_t *r_ptr, const int32_t *r_col, const double *r_val,
                    pq_loops **out);

/* The spanning tree the fundamental loops are built on. Loop j is its closing
 * branch (coefficient +1, running end_a -> end_b) plus the tree path back from
 * end_b to end_a; parent/up_branch/up_sign give each node's branch to its
 * parent (up_sign +1 where that branch runs from the parent to the node), and
 * level_nodes lists the non-root nodes by depth 1, 2, ... (level_ptr,
 * n_levels + 1 offsets). Mirrors spanning_tree's (up, upb, ups, depth)
 * and the loop walk of MQS._loops (src/pcbqs/mqs.py:48-129, 592-637). */
typedef struct pq_loop_tree {
    int32_t n_nodes;
    const int32_t *parent;       /* (n_nodes) parent node, -1 at a root           */
    const int32_t *up_branch;    /* (n_nodes) branch to the parent, -1 at a root  */
    const float *up_sign;        /* (n_nodes) +1 or -1                            */
    int32_t n_levels;
    const int32_t *level_ptr;    /* (n_levels + 1)                                */
    const int32_t *level_nodes;  /* (level_ptr[n_levels])                         */
    const int32_t *closing;      /* (n_loops) closing branch, loops in device order */
    const int32_t *end_a;        /* (n_loops) its first end node                  */
    const int32_t *end_b;        /* (n_loops) its second end node                 */
} pq_loop_tree;

/* ---- host topology (no device use) ------------------------------------------------------ */

typedef struct pq_loop_basis pq_loop_basis;   /* a fundamental loop basis on the host */
typedef struct pq_csr pq_csr;                 /* a host CSR matrix (int64 indices, FP64 values) */

/* The fundamental loop basis of a branch graph (spanning_tree and MQS._loops,
 * src/pcbqs/mqs.py:48-129, 592-637): branches k = 0 .. n_branches-1 run from
 * ends[2k] to ends[2k+1]; is_source marks the ports (never in the tree);
 * branches below n_copper are the copper (RWG edges and via segments: a
 * breadth-first forest from a central root per island, neighbours in
 * ascending order, each node pair's lowest-numbered branch), the others join
 * islands the copper leaves apart. One loop per non-tree branch whose ends the
 * tree joins (an RWG branch, k < n_rwg, with both ends on one node carries
 * none when ideal_terminals). Loops are ordered free loops first, along a
 * Morton curve of their centroids (the mean of their branches' branch_pos,
 * n_branches x 2), then the port loops. Host work only; thread-safe across
 * handles. */
int pq_loop_basis_create(int64_t n_nodes, int64_t n_branches, const int64_t *ends, const uint8_t *is_source,
                         int64_t n_copper, int64_t n_rwg, const double *branch_pos, int ideal_terminals,
                         pq_loop_basis **out);
/* info: [0] loops, [1] nonzeros of C, [2] free loops, [3] port loops. */
int pq_loop_basis_sizes(const pq_loop_basis *basis, int64_t info[4]);
/* Copies out (each pointer optional): tree (n_branches, 0/1), parent, upb, ups
 * (+1 where the parent branch runs from the parent to the node), depth
 * (n_nodes each, -1 / 0 at a root), cotree (loops: the closing branch), C in
 * CSC (colptr loops + 1, rows and vals nonzeros; rows ascending, vals +-1),
 * positions (loops x 2, the centroids) and port_branch (port loops). */
int pq_loop_basis_copy(const pq_loop_basis *basis, uint8_t *tree, int64_t *parent, int64_t *upb, int64_t *ups,
                       int64_t *depth, int64_t *cotree, int64_t *colptr, int64_t *rows, double *vals,
                       double *positions, int64_t *port_branch);
int pq_loop_basis_destroy(pq_loop_basis *basis);

/* Sheet resistance of ne RWG edges in total edge currents (MQS._resistance,
 * src/pcbqs/mqs.py:478-515): per triangle t and pair (m, n) of its edges,
 * alpha_m alpha_n (area_t / 3) sum_q (P_q - v_m).(P_q - v_n) / (sigma thickness_t)
 * / (l_m l_n), alpha = l / (2 area) on T+ (tp) and -l / (2 area) on T- (tm),
 * v the free vertices (vp, vm: ne x 2), P_q the 3-point rule points
 * (points: nt x 3 x 2). Summed over the triangles an edge pair shares. */
int pq_sheet_resistance(int64_t ne, const int64_t *tp, const int64_t *tm, const double *length, const double *vp,
                        const double *vm, int64_t nt, const double *area, const double *points,
                        const double *thickness, double sigma, pq_csr **out);
/* Rl = C^T R C (MQS._assemble's resistive part, src/pcbqs/mqs.py:639-660) with
 * R = blockdiag(re, diag(rv), 0): re the sheet resistance of the first re
 * rows' branches, rv (nv) the via barrels that follow, no resistance on the
 * remaining (port) branches. Rows and columns ascending, exact zeros dropped. */
int pq_loop_resistance(const pq_loop_basis *basis, const pq_csr *re, int64_t nv, const double *rv, pq_csr **out);
/* info: [0] rows, [1] columns, [2] nonzeros. */
int pq_csr_sizes(const pq_csr *m, int64_t info[3]);
int pq_csr_copy(const pq_csr *m, int64_t *ptr, int64_t *idx, double *val);
int pq_csr_destroy(pq_csr *m);

/* pq_loops_create with C given by its tree: U = Lb C from prefix sums of Lb's
 * rows down the tree (U[k][j] = Lb[k][closing j] + S_k(end_a j) - S_k(end_b j)),
 * then Ll = C^T U the same way on U's rows; FP64 sums and storage (U and Ll),
 * the same Ll as pq_loops_create. Lb must be symmetric (both triangles stored). */
int pq_loops_create_tree(int device, int n_loops, int n_free, const pq_loop_tree *tree, const void *Lb_dev,
                         int lb_precision, int64_t ldb, int n_branch_rows, const int32_t *r_ptr,
                         const int32_t *r_col, const double *r_val, pq_loops **out);
int pq_loops_destroy(pq_loops *loops);

/* The device loop inductance matrix (n_loops rows, leading dimension ldl): the
 * FP64 matrix (precision 64) or its FP32 copy (precision 32). */
int pq_loops_matrix(pq_loops *loops, int precision, void **Ll_dev, int64_t *ldl);

/* P in 1..8 or 16 (pad other counts by repeating a port column).
 * Port impedances Zp (P,P complex, interleaved re/im, row-major) at angular
 * frequency omega: the free loops eliminated, X = -Z_ff^-1 Z_fs, by restarted
 * lockstep GMRES(restart) with all P columns sharing every operator product
 * (precond 0: none, 1: block-Jacobi on 256-loop blocks of the free-loop order,
 * 2: two-level: the coarse space of pq_loops_set_coarse, then block-Jacobi),
 * then Zp = Z_ss + Z_sf X. X_host (n_free, 2P) (optional) receives the free
 * loop currents, real parts first; X0_host (optional, same layout) is the
 * initial iterate (zero when null). PQ_ECONVERGE when the true relative
 * residual exceeds max(5 tol, 1e-13). Replaces direct_ports
 * (src/pcbqs/mqs_ports.py:52-69) solved as port_matrix's iterative route
 * (src/pcbqs/mqs_iterative.py:298-367, _lockstep_gmres :109-170) with the
 * acceptance rule of peec_iterative._gmres (src/pcbqs/peec_iterative.py:100-112). */
int pq_loops_solve_ports(pq_loops *loops, double omega, int P, const int32_t *port_cols, int precond, double tol,
                         int restart, int maxiter, double *Zp, double *X_host, const double *X0_host,
                         int64_t *iterations, double *residual);

/* The direct port solve at omega: the same outputs as pq_loops_solve_ports
 * (Zp, optional X_host, the correction count in *steps and the final largest
 * relative residual in *residual), with Z_ff = Rl + j omega Ll formed in FP32
 * on the device and factored as L D L^T without pivoting (both parts are
 * symmetric positive definite), then refined against the FP64 system (its
 * residual taken with the FP64 Ll) until the relative residual is at most tol
 * (the reference uses 1e-12), for at most max_steps corrections, stopping once
 * a correction no longer halves it. A zero or non-finite FP32 pivot, or a
 * residual then above max(tol, 1e-10), sends the frequency to an FP64 LU with
 * partial pivoting of Z_ff (rocSOLVER zgetrf/zgetrs, loaded on first use), as
 * MixedLU refactors in double; PQ_ECONVERGE only if that too leaves a residual
 * above max(tol, 1e-10). The factor is kept for the next call at the same
 * omega. P 1..8 or 16. Replaces the dense magnetic solve MQS._solve_dense
 * (src/pcbqs/mqs.py:743-808) with lu_factor_rows / lu_solve_rows and MixedLU
 * (src/pcbqs/kernels.py:125-181, 266-292: FP32 LAPACK LU refined to 1e-12 in at
 * most 12 steps, an FP64 LU where that stalls). */
int pq_loops_solve_direct(pq_loops *loops, double omega, int P, const int32_t *port_cols, double tol,
                          int max_steps, double *Zp, double *X_host, int64_t *steps, double *residual);

/* Device time in milliseconds of the factorisation (forming Z_ff and its
 * L D L^T) that the last pq_loops_solve_direct used. No counterpart in the
 * reference. */
int pq_loops_factor_time(pq_loops *loops, double *ms);
/* Form and factor Z_ff = Rl + j omega Ll at n_omega frequencies together (the
 * L D L^T kernels take the frequency as a grid dimension: one launch per panel
 * step for all of them); pq_loops_solve_direct at any of these frequencies then
 * refines on its factor instead of factoring. *ms: device time of the batch.
 * Same factors as one pq_loops_solve_direct per frequency (the reference
 * factors per frequency, MQS._solve_dense, src/pcbqs/mqs.py:743-808). */
int pq_loops_factor_batch(pq_loops *loops, int n_omega, const double *omegas, double *ms);
/* pq_loops_solve_direct at n_omega frequencies at once (no currents): the
 * factorisations batched, the refinement in lockstep with every frequency's
 * solves in the same launches and each frequency's own stopping rule, so the
 * results are those of n_omega separate calls. Zp (n_omega, P, P) interleaved
 * complex, steps and residual (n_omega). */
int pq_loops_solve_direct_batch(pq_loops *loops, int n_omega, const double *omegas, int P, const int32_t *port_cols,
                                double tol, int max_steps, double *Zp, int64_t *steps, double *residual);

/* Direct port solves of `count` boards (loop handles on one device; board b
 * with P[b] ports, 1..8 or 16, at port_cols[b]) at the same n_omega
 * frequencies, each group of boards in ONE captured graph launch: boards taken
 * largest first and grouped while their free-loop counts stay within 3/4 of
 * the group's largest (each matrix padded with the identity to the group's
 * size), every board x frequency matrix formed, factored (FP32 L D L^T) and
 * solved in the same launches, refined against its board's FP64 system for up
 * to 2 steps with its own stopping rule kept on the device, and reduced to
 * port impedances. A matrix left above max(tol, 1e-10) is solved again alone
 * by pq_loops_solve_direct (further corrections, the FP64 fallback). Zp: per
 * board in order, (n_omega, P[b], P[b]) interleaved complex; steps and residual
 * (count, n_omega); stats[0] the graph launches, stats[1] the matrices solved
 * again alone. Replaces the per-board, per-frequency dense solves of the
 * reference's sweep (src/pcbqs/mqs_iterative.py:298-367 with the direct route
 * of port_method, :260-272). */
int pq_loops_solve_direct_multi(int count, pq_loops *const *handles, const int *P, const int32_t *const *port_cols,
                                int n_omega, const double *omegas, double tol, int max_steps, double *Zp,
                                int64_t *steps, double *residual, int64_t *stats);

/* The charge (PEEC) solve's electrostatic block: E = T^T P T / pscale for the
 * tree branches (end nodes ta, tb, nt of them; T's columns are their incidence
 * e_a - e_b), P_dev the node potential matrix (n_nodes square FP64, row stride
 * ldp), into E_dev (nt x nt FP64, row stride lde). Replaces
 * qform(T, P, T) / pscale of peec.prepare (src/pcbqs/peec.py:178-180). */
int pq_peec_electro(int device, const double *P_dev, int64_t ldp, int n_nodes, int nt, const int32_t *ta,
                    const int32_t *tb, double pscale, double *E_dev, int64_t lde);

/* The charge solve at angular frequency omega > 0 on a handle whose n = n_free
 * unknowns are nl loop currents then n - nl tree coordinates (its Ll the FP64
 * K^T Lb K, its Rl the sparse K^T R K): Z = Rl + j omega Ll, zscale = max |Z|,
 * S = Z / zscale on the first nl columns and j omega Z / pscale on the rest,
 * E_dev (n - nl square, row stride lde) added to the tree block; S X = B / zscale
 * by an FP64 LU with partial pivoting (rocSOLVER). B, X (n, nrhs) complex host,
 * interleaved, row-major; *zscale returned. At most 40000 unknowns. Replaces
 * peec_system with lu_factor_rows / lu_solve_rows of peec.solve
 * (src/pcbqs/kernels.py:63-78, src/pcbqs/peec.py:236-265). */
int pq_peec_solve(pq_loops *loops, int nl, const double *E_dev, int64_t lde, double pscale, double omega, int nrhs,
                  const double *B_host, double *X_host, double *zscale);

/* The free-loop right-hand sides B = -Z_fs (n_free, 2P) FP64 host of the port
 * loop columns at omega (the system pq_loops_solve_ports solves). Replaces the
 * right-hand side of direct_ports (src/pcbqs/mqs_ports.py:52-69). */
int pq_loops_rhs(pq_loops *loops, double omega, int P, const int32_t *port_cols, double *B_host);

/* W = Z_ff X = (Rl + j omega Ll) X on the free loops, X and W (n_free, 2P) FP64
 * host in the layout of pq_loops_rhs (real parts in columns 0..P-1, imaginary
 * parts in P..2P-1), P 1..8 or 16. precision 32: the GMRES matvec with the
 * FP32 Ll and the kernel the autotuner picked for 2P columns; 64: the FP64
 * system's product the residuals use. Replaces the product with A in
 * _lockstep_gmres (src/pcbqs/mqs_iterative.py:109-170). */
int pq_loops_operator(pq_loops *loops, double omega, int P, int precision, const double *X_host, double *W_host);

/* The coarse-space solution Phi (Phi^T Z_ff Phi)^-1 Phi^T B (n_free, 2P) FP64
 * host for the port right-hand sides at omega: the input of the learned
 * correction that refines it into an initial iterate. No counterpart in the
 * reference. */
int pq_loops_coarse_solution(pq_loops *loops, double omega, int P, const int32_t *port_cols, double *Xc_host);

/* The coarse space of the two-level preconditioner (precond 2 of
 * pq_loops_solve_ports): Phi (n_free, k) FP64 host, row-major (kept in FP32 on
 * the device). Each solve then applies y = Phi c + BJ(x - A Phi c),
 * c = (Phi^T A Phi)^-1 Phi^T x (coarse correction first, block-Jacobi on the
 * remainder). diagonal != 0 declares that Phi diagonalises both Galerkin
 * matrices Phi^T Rl Phi and Phi^T Ll Phi (a generalised eigenbasis), so each
 * frequency inverts only their diagonals. k = 0 removes it. No counterpart in
 * the reference (its iterative path uses a graph-Laplacian preconditioner,
 * src/pcbqs/mqs_precond.py:27-59). */
int pq_loops_set_coarse(pq_loops *loops, int k, const double *Phi, int diagonal);

/* pq_loops_set_coarse of the span of Phi (n_free, k) FP64 host, re-expressed on
 * the device in the generalised eigenbasis of its Galerkin pair (Rc, Lc) =
 * (Phi^T Rl Phi, Phi^T Ll Phi), FP64 (Ll the FP64 loop matrix). Where Rc's
 * Cholesky factor Rc = U^T U exists and shows lambda_min(Rc) > drop *
 * lambda_max(Rc) (1 / ||U^-1||_F^2 > drop * trace Rc), one eigensolve:
 * U^-T Lc U^-1 = Q Lam Q^T, Psi = Phi U^-1 Q; otherwise Rc = V S V^T, directions
 * with S below drop * max S removed, and (S^-1/2 V^T) Lc (V S^-1/2) = Q Lam Q^T,
 * Psi = Phi V S^-1/2 Q (rocSOLVER dpotrf / dsyevd). Either way Psi^T Rl Psi = I
 * and Psi^T Ll Psi = Lam, so each frequency's coarse solve is a diagonal
 * scaling. *k_out: the kept dimension. No counterpart in the reference. */
int pq_loops_set_coarse_diagonal(pq_loops *loops, int k, const double *Phi, double drop, int *k_out);

/* pq_loops_set_coarse_diagonal with Phi given as k sparse columns (CSC over the
 * n_free rows: colptr[k + 1], rows and vals[colptr[k]]), expanded on the device:
 * only the nonzeros cross PCIe. *path_out: 1 when the Cholesky route was taken,
 * 2 for the eigenvector route. No counterpart in the reference. */
int pq_loops_set_coarse_csc(pq_loops *loops, int k, const int32_t *colptr, const int32_t *rows,
                            const double *vals, double drop, int *k_out, int *path_out);

/* The block-Jacobi preconditioner's per-handle set-up (otherwise done by the
 * first solve that uses it): the generalised eigenvectors of every 256-loop
 * diagonal block's pencil (Ll_bb, Rl_bb), FP64 (rocSOLVER dsygvd, batched),
 * kept in FP32, so that each frequency's block inverses are
 * V diag(1 / (1 + j w lam)) V^T, one batched complex GEMM. *modal_out: 1 when
 * the modes are ready, -1 when a block's Rl_bb is not positive definite (then
 * every frequency inverts its blocks by Gauss-Jordan with partial pivoting),
 * 0 when disabled (PCBQS_AMD_BJ_MODAL=0). No counterpart in the reference. */
int pq_loops_prepare_block_jacobi(pq_loops *loops, int *modal_out);

/* The Galerkin matrices Phi^T Rl Phi and Phi^T Ll Phi (k x k FP64, row-major)
 * of the current coarse space, for a host-side change of basis. No counterpart
 * in the reference. */
int pq_loops_coarse_matrices(pq_loops *loops, double *Rc, double *Lc);

/* Capture each Arnoldi step's kernels once as a graph and replay it (enable 1),
 * launch them one by one (0), or follow PCBQS_AMD_GRAPHS (-1, the default:
 * on when it is "1"). A capture is invalidated by another thread's
 * synchronising runtime calls, so several handles driven from several threads
 * at once should run without graphs. No counterpart in the reference. */
int pq_loops_use_graphs(pq_loops *loops, int enable);

/* The operator product's autotuned kernel for NR real columns (tuned on first
 * use: candidates timed with HIP events on this handle's matrix): chunk 0 is
 * the symmetric kernel that reads the tiles on and above the diagonal of Ll
 * only (NR <= 16), chunk > 0 the column-chunk width of the whole-matrix panel
 * kernel; and its measured time per product in microseconds (0 when not tuned:
 * fewer than 1024 free loops). PCBQS_AMD_OPERATOR=panel or =symmetric pins the
 * choice. No counterpart in the reference. */
int pq_loops_tuning(pq_loops *loops, int NR, int *chunk, double *microseconds);

/* Statistics of the last solve: iterations, operator products, relative
 * residual, wall seconds. Not in the reference (solve_info's role,
 * src/pcbqs/mqs_iterative.py:229-230). */
int pq_loops_last_solve(pq_loops *loops, int64_t *iterations, int64_t *matvecs, double *residual, double *seconds);
/* The number of frequencies the last direct solve (pq_loops_solve_direct or
 * pq_loops_solve_direct_batch) solved with the FP64 fallback factor (the role of
 * MixedLU.double, src/pcbqs/kernels.py:147-154, 179-181). */
int pq_loops_last_fallbacks(pq_loops *loops, int64_t *count);

/* ---- compressed (H-matrix) operator product ------------------------------ */

typedef struct pq_hmatrix pq_hmatrix;   /* a block-compressed operator on one device */

/* Upload the blocks of a transpose-symmetric compressed operator of order n.
 * Block q (meta[6q .. 6q+5] = a_off, la, b_off, lb, rank, same) holds rows
 * idx[a_off .. a_off+la) and columns idx[b_off .. b_off+lb) of A: dense
 * (rank -1, U la x lb) or low rank (rank 0..256, U la x rank then V rank x lb);
 * unless `same`, it also stands for its transpose at (b, a). `factors` is the
 * blocks' U (and V) row-major one after another, FP64 or (is_complex)
 * interleaved complex FP64, n_factor elements. Replaces the packed layout of
 * HMatrix._pack / HBlocks (src/pcbqs/hmatrix.py:235-270,
 * src/pcbqs/core/hmatrix_kernels.hpp:103-160): inconsistent layouts return
 * PQ_EINVAL. */
int pq_hmatrix_create(int device, int64_t n, int64_t n_blocks, const int64_t *meta, const int32_t *idx,
                      int64_t n_idx, const void *factors, int64_t n_factor, int is_complex, pq_hmatrix **out);
int pq_hmatrix_destroy(pq_hmatrix *h);
/* info[0..4] = n, contribution slots, factor bytes, dense blocks, low-rank blocks. */
int pq_hmatrix_info(pq_hmatrix *h, int64_t info[5]);

/* Y = A X, X and Y (n x k) row-major of the operator's element type, host
 * memory, synchronised; deterministic (every output row sums its blocks'
 * contributions in block order). Replaces HMatrix._matmat /
 * HBlocks.apply / hmatrix_product (src/pcbqs/hmatrix.py:223-297,
 * src/pcbqs/core/hmatrix_kernels.hpp:20-101). */
int pq_hmatrix_apply(pq_hmatrix *h, int64_t k, const void *x, void *y);
/* The same on device buffers, asynchronous on `stream` (NULL: the handle's). */
int pq_hmatrix_apply_device(pq_hmatrix *h, int64_t k, const void *x_dev, void *y_dev, void *stream);
/* Mean device milliseconds of `reps` products on the vectors of the last
 * pq_hmatrix_apply (HIP events). Measurement only: no counterpart in the reference. */
int pq_hmatrix_time(pq_hmatrix *h, int reps, double *ms);

/* ---- construction of the compressed RWG inductance on the device --------------
 * An entry provider over a prepared fill geometry: ne interior edges with
 * support (ne, 2) plus/minus triangles, vertices (ne, 2, 2) their free
 * vertices and alpha (ne, 2) = (1/(2 a+), -1/(2 a-)), in the geometry's
 * (centred) frame; entries are scale * sum of the four support pairs'
 * canonical moments (near rule `near`, `outer_levels`). Replaces
 * RwgEntries (src/pcbqs/core/hmatrix_build.hpp:27-43) and the provider
 * set-up of edge_inductance (src/pcbqs/compressed.py:61-73). The provider
 * also holds every block built through it until it is destroyed. */
typedef struct pq_entries pq_entries;
int pq_rwg_entries_create(pq_fill *f, int64_t ne, const int32_t *support, const double *vertices,
                          const double *alpha, double scale, double near, int outer_levels, pq_entries **out);
int pq_entries_destroy(pq_entries *e);
/* out (nr, nc) row-major FP64 host = entries (rows[i], cols[j]). Replaces
 * rwg_inductance (src/pcbqs/core/compressed_kernels.hpp:127-144). */
int pq_entries_block(pq_entries *e, const int32_t *rows, int64_t nr, const int32_t *cols, int64_t nc, double *out);
/* Checked ACA of nblocks admissible blocks (rows[row_ptr[q]:row_ptr[q+1]] x
 * cols[col_ptr[q]:col_ptr[q+1]]) to relative tolerance tol, all on the
 * device. ids[q] = the kept block's id (>= 0) and ranks[q] its rank for a
 * block that converged, -1 for one that did not (the reference's None);
 * *entries (optional) = entries evaluated. Factors stay on the device.
 * Replaces aca_blocks / aca (src/pcbqs/core/hmatrix_build.hpp:143-327). */
int pq_aca_blocks(pq_entries *e, int64_t nblocks, const int64_t *row_ptr, const int32_t *rows,
                  const int64_t *col_ptr, const int32_t *cols, double tol, int64_t *ids, int64_t *ranks,
                  int64_t *entries);
/* Dense leaves, same[q] != 0 marking a diagonal block (symmetrised,
 * (B + B^T) / 2). ids[q] = the kept block's id. Replaces dense_blocks
 * (hmatrix_build.hpp:329-352) and the symmetrisation of hmatrix.py:204-206. */
int pq_dense_blocks(pq_entries *e, int64_t nblocks, const int64_t *row_ptr, const int32_t *rows,
                    const int64_t *col_ptr, const int32_t *cols, const int32_t *same, int64_t *ids);
/* Host copies of kept block `id`: dense (m x n) into u, or U (m x r) and V
 * (r x n), row-major. */
int pq_entries_factors(pq_entries *e, int64_t id, double *u, double *v);
/* The product handle of the kept blocks ids[0..nblocks) in that order (the
 * HMatrix block list, hmatrix.py:80-83), factors gathered on the device. */
int pq_entries_hmatrix(pq_entries *e, int64_t nblocks, const int64_t *ids, pq_hmatrix **out);
/* Entries evaluated, ACA blocks run again with the worst-case check room,
 * factor bytes held, and device milliseconds of the ACA and dense launches
 * (ms[0], ms[1]). Measurement only: no counterpart in the reference. */
int pq_entries_stats(pq_entries *e, int64_t *entries, int64_t *reruns, int64_t *bytes, double *ms);

/* ---- Biot-Savart field --------------------------------------------------------
 * H (np, 3, K, 2) FP64 host (real, imaginary) at points pts (np, 3) of ns sheet
 * sources src (ns, 3 + 4K: position, then Jx re, Jx im, Jy re, Jy im per
 * excitation, each current times its quadrature weight) and nseg straight
 * filaments seg (nseg, 6 + 2K: p0, p1, then I re, I im per excitation), K in
 * {1, 2, 4, 8}. precision 64 (parity) or 32 computes on the device; kernel_ms
 * (optional) receives the kernels' time. Replaces _h_tris and h_segments
 * (src/pcbqs/biot.py:24-51, 82-105); the caller builds the refined source list
 * of h_sheets (biot.py:54-79). */
int pq_biot_field(int device, int precision, int K, int64_t ns, const double *src, int64_t nseg, const double *seg,
                  int64_t np, const double *pts, double *H, double *kernel_ms);

/* The near field on a grid: nlayers current densities J (layer, [Jx, Jy],
 * sx, sy, K) interleaved complex FP64 on a uniform source grid of spacing dx,
 * dy (host deposit), layer heights layer_z, convolved with the Biot-Savart
 * kernel of each height difference by a linear fx x fy FFT on the device
 * (fx >= 3 sx - 2, fy >= 3 sy - 2; hipFFT, FP64), summed over layers; H
 * (nox, noy, nh, 3, K) interleaved complex = the full convolution's points
 * (ox0 + step i, oy0 + step j). *kernel_ms: device time. Replaces the FFT
 * part of h_grid (src/pcbqs/biot.py:108-181). */
int pq_biot_grid(int device, int nlayers, const double *layer_z, int K, int sx, int sy, const double *J, double dx,
                 double dy, int nh, const double *heights, int fx, int fy, int ox0, int oy0, int step, int nox,
                 int noy, double *H, double *kernel_ms);

/* ---- message-passing network (the learned coarse-space correction) -------
 * No counterpart in the reference. All pointers are device pointers; stream is
 * a hipStream_t (NULL: the default stream). Activations are FP32 row-major;
 * the dense products round both operands to BF16 and accumulate in FP32 on
 * the matrix cores. Supported (k1, k2, m) shapes: (16, 0, 64), (64, 64, 64),
 * (64, 0, 128), (64, 0, 16). */

/* C (N, m) = act([A1 | A2] W + bias) (+ res): A1 (N, k1), A2 (N, k2), W
 * (k1 + k2, m) row-major; relu != 0 applies max(0, .) before the residual res
 * (optional); Zpre (optional, (N, m)) receives the pre-activation. */
int pq_mpnn_gemm(int N, const float *A1, int lda1, int k1, const float *A2, int lda2, int k2, const float *W, int m,
                 const float *bias, const float *res, int ldres, int relu, float *C, int ldc, float *Zpre,
                 void *stream);

/* Workspace (floats) that pq_mpnn_gemm_tn needs for N rows and a (k, m) weight. */
int pq_mpnn_workspace(int N, int k, int m, size_t *floats);

/* Weight gradients: dW (k1 + k2, m) = [A1 | A2]^T G and db (m) = column sums
 * of G (N, m), summed in a fixed order. */
int pq_mpnn_gemm_tn(int N, const float *A1, int lda1, int k1, const float *A2, int lda2, int k2, const float *G,
                    int ldg, int m, float *work, size_t work_floats, float *dW, float *db, void *stream);

/* out (N, 64) = base + base2 + (sum over the CSR row i of w[t] in[col[t]]):
 * the neighbour aggregation (base, base2 optional). */
int pq_mpnn_aggregate(int N, const int32_t *ptr, const int32_t *col, const float *w, const float *in, int ldi,
                      const float *base, int ldb, const float *base2, int ldb2, float *out, int ldo, void *stream);

/* out = g where z > 0 else 0 (count elements): the ReLU's backward. */
int pq_mpnn_relu_mask(size_t count, const float *z, const float *g, float *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
