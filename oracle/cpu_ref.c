/*
 * oracle/cpu_ref.c -- TEST INFRASTRUCTURE ONLY (the parity checker and the timed
 * CPU baseline). Nothing in the product loads this library.
 *
 * A plain-C restatement of the reference's FP64 triangle-pair fill:
 *   - cr_pair          <- src/pcbqs/core/static_kernels.hpp:53-106  (static_pair)
 *   - cr_pair_kind     <- src/pcbqs/core/static_kernels.hpp:130-136 (pair_kind),
 *                         src/pcbqs/static_backends.py:288-298       (FillEngine.kinds)
 *   - cr_fill          <- src/pcbqs/core/static_kernels.hpp:184-252 (static_fill)
 *   - cr_pairs         <- src/pcbqs/core/static_kernels.hpp:147-172 (static_pairs)
 *   - cr_combine       <- src/pcbqs/core/static_kernels.hpp:257-310 (static_combine)
 *   - cr_project       <- src/pcbqs/mqs.py:665-667 (_project_inductance: qform(C, L, C),
 *                         C^T L C of the sparse loop basis and the dense branch matrix)
 * The arithmetic is evaluated in the same order as the reference so that, built
 * with -ffp-contract=off, results agree with the reference core bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* Read-only triangle geometry, all float64, C order (the layout of
 * FillGeometry: src/pcbqs/static_backends.py:105-130). */
typedef struct {
    int32_t nt;
    const double *V;      /* (nt,3,2) vertices (centred)            */
    const double *L;      /* (nt,3)   edge lengths                  */
    const double *tau;    /* (nt,3,2) unit edge tangents            */
    const double *un;     /* (nt,3,2) outward edge normals          */
    const double *area;   /* (nt)                                   */
    const double *P3;     /* (nt,3,2) far-rule points (_BAR3)       */
    const double *qp[3];  /* (nt,nq[k],2) observation points, level k */
    const double *qw[3];  /* (nq[k]) observation weights            */
    int32_t nq[3];
    const double *z;      /* (nt) sheet heights                     */
    const double *cent;   /* (nt,2) centroids                       */
    const double *radius; /* (nt) bounding radii                    */
} cr_geometry;

/* Six moments of the ordered pair (observation p, source q) at vertical offset d:
 * m = [M0, Mro_x, Mro_y, Mrs_x, Mrs_y, Mrr]. k < 0: 3x3 product rule; k >= 0:
 * closed-form source integral over q at the level-k observation points of p. */
void cr_pair(const cr_geometry *g, int p, int q, double d, int k, int scalar_only, double *m)
{
    const double ad = fabs(d);
    for (int j = 0; j < 6; ++j) m[j] = 0.0;
    if (k < 0) {
        const double w = g->area[p] * g->area[q] / 9;
        const double *op = g->P3 + 6 * (size_t)p, *sq = g->P3 + 6 * (size_t)q;
        for (int a = 0; a < 3; ++a) {
            const double x = op[2 * a], y = op[2 * a + 1];
            for (int b = 0; b < 3; ++b) {
                const double sx = sq[2 * b], sy = sq[2 * b + 1];
                const double dx = x - sx, dy = y - sy;
                const double gg = w / sqrt(dx * dx + dy * dy + d * d);
                m[0] += gg;
                if (!scalar_only) {
                    m[1] += gg * x; m[2] += gg * y; m[3] += gg * sx; m[4] += gg * sy;
                    m[5] += gg * (x * sx + y * sy);
                }
            }
        }
        return;
    }
    const int nq = g->nq[k];
    const double *pts = g->qp[k] + 2 * (size_t)p * nq, *wts = g->qw[k];
    const double *Vq = g->V + 6 * (size_t)q, *Tq = g->tau + 6 * (size_t)q;
    const double *Nq = g->un + 6 * (size_t)q, *Lq = g->L + 3 * (size_t)q;
    for (int j = 0; j < nq; ++j) {
        const double x = pts[2 * j], y = pts[2 * j + 1];
        double I1 = 0, IRx = 0, IRy = 0;
        for (int e = 0; e < 3; ++e) {
            const double vx = Vq[2 * e] - x, vy = Vq[2 * e + 1] - y;
            const double sm = vx * Tq[2 * e] + vy * Tq[2 * e + 1];
            const double sp = sm + Lq[e];
            const double t0 = vx * Nq[2 * e] + vy * Nq[2 * e + 1];
            const double r02 = t0 * t0 + d * d, r0 = sqrt(r02);
            const double rm = sqrt(r02 + sm * sm), rp = sqrt(r02 + sp * sp);
            const double lg = r0 > 1e-14 * Lq[e] ? asinh(sp / r0) - asinh(sm / r0) : 0;
            double term = t0 * lg;
            if (ad > 0) term -= ad * (atan2(t0 * sp, r02 + ad * rp) - atan2(t0 * sm, r02 + ad * rm));
            I1 += term;
            if (!scalar_only) {
                const double er = .5 * (r02 * lg + sp * rp - sm * rm);
                IRx += er * Nq[2 * e]; IRy += er * Nq[2 * e + 1];
            }
        }
        const double w = g->area[p] * wts[j];
        m[0] += w * I1;
        if (!scalar_only) {
            const double sx = IRx + x * I1, sy = IRy + y * I1;
            m[1] += w * x * I1; m[2] += w * y * I1;
            m[3] += w * sx; m[4] += w * sy; m[5] += w * (x * sx + y * sy);
        }
    }
}

/* The near/far rule: -1 far, `outer` when the centroids are within the summed
 * radii, 0 otherwise; near = +inf makes every pair analytic. Built without
 * contraction, every product here is rounded before it is added. */
int cr_pair_kind(const cr_geometry *g, int p, int q, double d, double near, int outer)
{
    const double dx = g->cent[2 * p] - g->cent[2 * q], dy = g->cent[2 * p + 1] - g->cent[2 * q + 1];
    const double dist = sqrt((dx * dx + dy * dy) + d * d);
    const double size = g->radius[p] + g->radius[q];
    if (!isinf(near) && !(dist <= near * size)) return -1;
    return dist <= size ? outer : 0;
}

/* Explicit pair list: out (n, 6) or (n, 1). Returns 0, or -1 on invalid input. */
int cr_pairs(const cr_geometry *g, const int32_t *P, const int32_t *Q, const double *D,
             const int32_t *K, int64_t n, int scalar_only, int workers, double *out)
{
    const int width = scalar_only ? 1 : 6;
    for (int64_t i = 0; i < n; ++i)
        if (P[i] < 0 || P[i] >= g->nt || Q[i] < 0 || Q[i] >= g->nt || K[i] < -1 || K[i] > 2 || !isfinite(D[i]))
            return -1;
#pragma omp parallel for num_threads(workers) schedule(guided) if (workers > 1)
    for (int64_t i = 0; i < n; ++i) {
        double m[6];
        cr_pair(g, P[i], Q[i], D[i], K[i], scalar_only, m);
        for (int j = 0; j < width; ++j) out[i * width + j] = m[j];
    }
    return 0;
}

/* Whole reciprocal fill into caller-owned row-major (nt, nt) planes
 * M[0..3] = M0, Mrx, Mry, Mrr (M[0] only when scalar_only). Unordered pairs
 * p <= q in 64 x 64 tiles; each pair's owner writes both reciprocal entries.
 * Returns 0 (ok), 1 (non-finite moments) or -1 (bad input); counts pairs. */
int cr_fill(const cr_geometry *g, double near, int outer, int scalar_only, int workers,
            double *const *M, int64_t *pairs, int64_t *analytic_out)
{
    const int nt = g->nt, side = 64, blocks = (nt + side - 1) / side;
    if (workers < 1 || outer < 0 || outer > 2 || isnan(near) || near < 0) return -1;
    const long ntiles = (long)blocks * (blocks + 1) / 2;
    int *tp = malloc(sizeof(int) * (size_t)(ntiles ? ntiles : 1)), *tq = malloc(sizeof(int) * (size_t)(ntiles ? ntiles : 1));
    long t = 0;
    for (int bp = 0; bp < blocks; ++bp)
        for (int bq = bp; bq < blocks; ++bq) { tp[t] = bp; tq[t] = bq; ++t; }
    int64_t analytic = 0;
    int finite = 1;
    const int width = scalar_only ? 1 : 6;
#pragma omp parallel num_threads(workers) if (workers > 1) reduction(+ : analytic) reduction(&& : finite)
    {
        double *buf = malloc(sizeof(double) * (size_t)side * side * 6);
#pragma omp for schedule(dynamic, 1)
        for (long tile = 0; tile < ntiles; ++tile) {
            const int p0 = tp[tile] * side, q0 = tq[tile] * side;
            const int p1 = p0 + side < nt ? p0 + side : nt, q1 = q0 + side < nt ? q0 + side : nt;
            const int diag = p0 == q0;
            for (int p = p0; p < p1; ++p)
                for (int q = diag ? p : q0; q < q1; ++q) {
                    const double d = g->z[p] - g->z[q];
                    const int k = cr_pair_kind(g, p, q, d, near, outer);
                    analytic += k >= 0;
                    double *m = buf + 6 * ((size_t)(p - p0) * side + (q - q0));
                    cr_pair(g, p, q, d, k, scalar_only, m);
                    for (int j = 0; j < width; ++j) finite = finite && isfinite(m[j]);
                }
            for (int p = p0; p < p1; ++p)
                for (int q = diag ? p : q0; q < q1; ++q) {
                    const double *m = buf + 6 * ((size_t)(p - p0) * side + (q - q0));
                    const size_t pq = (size_t)p * nt + q;
                    M[0][pq] = m[0];
                    if (!scalar_only) {
                        M[1][pq] = p == q ? (m[1] + m[3]) / 2 : m[1];
                        M[2][pq] = p == q ? (m[2] + m[4]) / 2 : m[2];
                        M[3][pq] = m[5];
                    }
                }
            for (int q = q0; q < q1; ++q)
                for (int p = p0; p < (diag ? (q < p1 ? q : p1) : p1); ++p) {
                    const double *m = buf + 6 * ((size_t)(p - p0) * side + (q - q0));
                    const size_t qp = (size_t)q * nt + p;
                    M[0][qp] = m[0];
                    if (!scalar_only) { M[1][qp] = m[3]; M[2][qp] = m[4]; M[3][qp] = m[5]; }
                }
        }
        free(buf);
    }
    free(tp); free(tq);
    if (pairs) *pairs = (int64_t)nt * (nt + 1) / 2;
    if (analytic_out) *analytic_out = analytic;
    return finite ? 0 : 1;
}

/* RWG Galerkin combination (symmetric), out (ne, ne):
 * out[m,n] = f s_m s_n sum_{i,j} a_mi a_nj [Mrr(p,q) - v_nj.Mr(p,q) - v_mi.Mr(q,p) + v_mi.v_nj M0(p,q)]
 * with p = tri[i][m], q = tri[j][n]; tri (2, ne) int32, alpha (2, ne), vertex (2, ne, 2). */
int cr_combine(const double *Mrr, const double *Mrx, const double *Mry, const double *M0, int64_t nt,
               const int32_t *tri, const double *alpha, const double *vertex, const double *scale,
               int64_t ne, double factor, int workers, double *out)
{
    for (int64_t i = 0; i < 2 * ne; ++i)
        if (tri[i] < 0 || tri[i] >= nt) return -1;
    const int64_t side = 32, blocks = (ne + side - 1) / side, ntiles = blocks * (blocks + 1) / 2;
#pragma omp parallel for num_threads(workers) schedule(dynamic, 1) if (workers > 1)
    for (int64_t tile = 0; tile < ntiles; ++tile) {
        int64_t bm = 0, rest = tile;
        while (rest >= blocks - bm) { rest -= blocks - bm; ++bm; }
        const int64_t bn = bm + rest;
        const int64_t m0 = bm * side, m1 = m0 + side < ne ? m0 + side : ne;
        const int64_t n0 = bn * side, n1 = n0 + side < ne ? n0 + side : ne;
        for (int64_t m = m0; m < m1; ++m)
            for (int64_t n = bm == bn ? m : n0; n < n1; ++n) {
                double sum = 0;
                for (int i = 0; i < 2; ++i) {
                    const size_t p = (size_t)tri[i * ne + m];
                    const double am = alpha[i * ne + m], vmx = vertex[2 * (i * ne + m)], vmy = vertex[2 * (i * ne + m) + 1];
                    for (int j = 0; j < 2; ++j) {
                        const size_t q = (size_t)tri[j * ne + n];
                        const double vnx = vertex[2 * (j * ne + n)], vny = vertex[2 * (j * ne + n) + 1];
                        const size_t pq = p * nt + q, qp = q * nt + p;
                        sum += am * alpha[j * ne + n] * (Mrr[pq] - vnx * Mrx[pq] - vny * Mry[pq] - vmx * Mrx[qp] -
                                                         vmy * Mry[qp] + (vmx * vnx + vmy * vny) * M0[pq]);
                    }
                }
                const double v = factor * scale[m] * scale[n] * sum;
                out[m * ne + n] = v;
                out[n * ne + m] = v;
            }
    }
    return 0;
}

/* Ll (nl, nl) = C^T L C for the sparse loop basis C (n rows; CSC: colptr (nl + 1), rows, sign)
 * and the symmetric dense L (n, n, row stride ld). First Ut = (L C)^T, row j the signed sum of
 * the rows of L on loop j's branches (L's rows stand for its columns), then Ll[j][i] = the signed
 * sum of Ut[j] over loop i's branches. Ut (nl, n) is the caller's scratch. */
int cr_project(const double *L, int64_t ld, int64_t n, const int32_t *colptr, const int32_t *rows,
               const double *sign, int64_t nl, int workers, double *Ut, double *Ll)
{
    for (int64_t j = 0; j < nl; ++j)
        for (int32_t t = colptr[j]; t < colptr[j + 1]; ++t)
            if (rows[t] < 0 || rows[t] >= n) return -1;
#pragma omp parallel for num_threads(workers) schedule(dynamic, 4) if (workers > 1)
    for (int64_t j = 0; j < nl; ++j) {
        double *u = Ut + (size_t)j * n;
        for (int64_t k = 0; k < n; ++k) u[k] = 0.0;
        for (int32_t t = colptr[j]; t < colptr[j + 1]; ++t) {
            const double s = sign[t], *row = L + (size_t)rows[t] * ld;
            for (int64_t k = 0; k < n; ++k) u[k] += s * row[k];
        }
    }
#pragma omp parallel for num_threads(workers) schedule(dynamic, 4) if (workers > 1)
    for (int64_t j = 0; j < nl; ++j) {
        const double *u = Ut + (size_t)j * n;
        for (int64_t i = 0; i < nl; ++i) {
            double acc = 0.0;
            for (int32_t t = colptr[i]; t < colptr[i + 1]; ++t) acc += sign[t] * u[rows[t]];
            Ll[(size_t)j * nl + i] = acc;
        }
    }
    return 0;
}

int cr_max_threads(void) { return omp_get_max_threads(); }
