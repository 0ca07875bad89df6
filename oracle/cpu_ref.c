/*
 * oracle/cpu_ref.c -- CPU restatement of CyLag's particle time step.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline); see cpu_ref.h.
 * Every function cites the reference file:line it restates.  Build:
 *   gcc -O2 -ffp-contract=off -fno-fast-math -fPIC -shared  (oracle/Makefile)
 */
#include "cpu_ref.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* cylag/core/constants.pyx:4-75 */
#define EPSILON 1.e-16
#define EPSILON_LSM2 1.e-8
#define EPSILON_REBOUND 1.e-8
#define DEG_TO_RAD 0.017453292519943295
#define TWO_SQRT3 3.4641016151377544
#define BND_WALL_REF 0
#define BND_OPEN_REF 1
#define MAX_VERTEX_ADJACENCY 15
#define PART_LOC_O (-1)
#define PART_LOC_B (-2)
#define PART_LOC_A (-3)
#define GRAV 9.80665
#define NU0 1.3e-6
#define C0KOLM 1.2
#define CMU 0.09
#define KAPPA 0.41

/* ========================================================================== */
/* PCG32 -- cylag/lsm/random_utils.pyx:12-67                                  */
/* ========================================================================== */
static inline uint32_t pcg_next(ref_pcg32 *g)
{   /* random_utils.pyx:22-28 */
    uint64_t old = g->state;
    g->state = old * 6364136223846793005ULL + g->inc;
    uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
    uint32_t rot = (uint32_t)(old >> 59);
    return (xs >> rot) | (xs << ((32 - rot) & 31));
}
uint32_t ref_pcg32_next(ref_pcg32 *g) { return pcg_next(g); }

void ref_pcg32_seed(ref_pcg32 *g, uint64_t seed, uint64_t stream)
{   /* random_utils.pyx:30-37 */
    g->state = 0;
    g->inc = (stream << 1) | 1u;
    pcg_next(g);
    g->state += seed;
    pcg_next(g);
}

void ref_pcg32_advance(ref_pcg32 *g, uint64_t delta)
{   /* random_utils.pyx:39-53 */
    uint64_t acc_mult = 1, acc_plus = 0;
    uint64_t cur_mult = 6364136223846793005ULL, cur_plus = g->inc;
    while (delta > 0) {
        if (delta & 1) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1) * cur_plus;
        cur_mult *= cur_mult;
        delta >>= 1;
    }
    g->state = acc_mult * g->state + acc_plus;
}

double ref_random_uniform(ref_pcg32 *g)
{   /* random_utils.pyx:55-67 */
    uint32_t u = pcg_next(g);
    return u * (1.0 / 4294967296.0);
}

double ref_random_gaussian(ref_pcg32 *g)
{   /* random_utils.pyx:72-88 (Marsaglia polar, single) */
    double x1 = 0., x2 = 0., w = 2.0;
    while (w >= 1.0 || w == 0.0) {
        x1 = 2.0 * ref_random_uniform(g) - 1.0;
        x2 = 2.0 * ref_random_uniform(g) - 1.0;
        w = x1 * x1 + x2 * x2;
    }
    w = sqrt((-2.0 * log(w)) / w);
    return x1 * w;
}

void ref_random_gaussian_pair(ref_pcg32 *g, double *a, double *b)
{   /* random_utils.pyx:90-108 */
    double x1 = 0., x2 = 0., w = 2.0;
    while (w >= 1.0 || w == 0.0) {
        x1 = 2.0 * ref_random_uniform(g) - 1.0;
        x2 = 2.0 * ref_random_uniform(g) - 1.0;
        w = x1 * x1 + x2 * x2;
    }
    w = sqrt((-2.0 * log(w)) / w);
    *a = x1 * w;
    *b = x2 * w;
}

/* Ziggurat tables, random_utils.pyx:113-149 */
static const double ZIG_R = 3.442619855899;
static uint32_t zkn[128];
static double zwn[128], zfn[128];
static int zig_ready = 0;

void ref_ziggurat_tables(uint32_t *kn, double *wn, double *fn)
{   /* random_utils.pyx:118-146 */
    int i;
    double m1 = 2147483648.0;
    double dn = 3.442619855899, tn = dn, vn = 9.91256303526217e-3, q;
    q = vn / exp(-0.5 * dn * dn);
    kn[0] = (uint32_t)((dn / q) * m1);
    kn[1] = 0;
    wn[0] = q / m1;
    wn[127] = dn / m1;
    fn[0] = 1.0;
    fn[127] = exp(-0.5 * dn * dn);
    for (i = 126; i > 0; i--) {
        dn = sqrt(-2.0 * log(vn / dn + exp(-0.5 * dn * dn)));
        kn[i + 1] = (uint32_t)((dn / tn) * m1);
        tn = dn;
        fn[i] = exp(-0.5 * dn * dn);
        wn[i] = dn / m1;
    }
}

static void zig_init(void)
{
    if (!zig_ready) { ref_ziggurat_tables(zkn, zwn, zfn); zig_ready = 1; }
}

double ref_random_gaussian_ziggurat(ref_pcg32 *g)
{   /* random_utils.pyx:154-194 */
    int32_t hz;
    uint32_t iz, ahz;
    double x, y;
    zig_init();
    for (;;) {
        hz = (int32_t)pcg_next(g);
        iz = (uint32_t)hz & 127u;
        /* libc abs(INT_MIN) wraps to INT_MIN: 0x80000000 as uint32 */
        ahz = hz < 0 ? (uint32_t)(-(int64_t)hz) : (uint32_t)hz;
        if (ahz < zkn[iz]) return hz * zwn[iz];
        if (iz == 0) {
            for (;;) {
                x = -log(ref_random_uniform(g)) * 0.2904764;
                y = -log(ref_random_uniform(g));
                if (y + y >= x * x) {
                    if (hz > 0) return ZIG_R + x;
                    return -ZIG_R - x;
                }
            }
        }
        x = hz * zwn[iz];
        if (zfn[iz] + ref_random_uniform(g) * (zfn[iz - 1] - zfn[iz]) < exp(-0.5 * x * x))
            return x;
    }
}

int ref_generate_random_pair(ref_pcg32 *g, int method, double *r1, double *r2)
{   /* random_utils.pyx:202-233 */
    if (method == 1) {
        *r1 = (ref_random_uniform(g) - 0.5) * TWO_SQRT3;
        *r2 = (ref_random_uniform(g) - 0.5) * TWO_SQRT3;
    } else if (method == 2) {
        *r1 = ref_random_gaussian_ziggurat(g);
        *r2 = ref_random_gaussian_ziggurat(g);
    } else if (method == 3) {
        ref_random_gaussian_pair(g, r1, r2);
    } else return REF_ERR_RNG_METHOD;
    return REF_OK;
}

int ref_generate_random_single(ref_pcg32 *g, int method, double *r)
{   /* random_utils.pyx:235-262 */
    if (method == 1) *r = (ref_random_uniform(g) - 0.5) * TWO_SQRT3;
    else if (method == 2) *r = ref_random_gaussian_ziggurat(g);
    else if (method == 3) *r = ref_random_gaussian(g);
    else return REF_ERR_RNG_METHOD;
    return REF_OK;
}

void ref_fill_uint32(ref_pcg32 *g, int64_t n, uint32_t *out)
{
    for (int64_t i = 0; i < n; i++) out[i] = pcg_next(g);
}

int ref_fill_normals(ref_pcg32 *g, int method, int64_t n, double *out)
{
    for (int64_t i = 0; i < n; i++) {
        int e = ref_generate_random_single(g, method, &out[i]);
        if (e) return e;
    }
    return REF_OK;
}

/* ========================================================================== */
/* Mesh tables -- cylag/geom/triangular_mesh.pyx                              */
/* ========================================================================== */
static int cmp_edge(const void *a, const void *b)
{
    const int *ea = (const int *)a, *eb = (const int *)b;
    if (ea[0] != eb[0]) return ea[0] < eb[0] ? -1 : 1;
    if (ea[1] != eb[1]) return ea[1] < eb[1] ? -1 : 1;
    return 0;
}

int ref_build_edges(int nt, const int *tri, int *e)
{   /* triangular_mesh.pyx:288-309: sides (min,max) then np.unique(axis=0) */
    int p, n = 3 * nt, ne = 0;
    for (p = 0; p < n; p++) {
        int k = p / 3, j = p % 3;
        int a = tri[3 * k + j], b = tri[3 * k + (j + 1) % 3];
        e[2 * p] = a < b ? a : b;
        e[2 * p + 1] = a < b ? b : a;
    }
    qsort(e, (size_t)n, 2 * sizeof(int), cmp_edge);
    for (p = 0; p < n; p++) {
        if (ne == 0 || e[2 * p] != e[2 * (ne - 1)] || e[2 * p + 1] != e[2 * (ne - 1) + 1]) {
            e[2 * ne] = e[2 * p];
            e[2 * ne + 1] = e[2 * p + 1];
            ne++;
        }
    }
    return ne;
}

int ref_build_connectivity(int nv, int nt, const int *tri, int ne, const int *edges,
                           int *edgetri, int *triedge)
{
    const int nn = MAX_VERTEX_ADJACENCY;
    int i, j, k, m, n;
    int *dyn = (int *)calloc((size_t)nv, sizeof(int));
    int *adj = (int *)malloc((size_t)nv * nn * sizeof(int));
    int trilist[2 * MAX_VERTEX_ADJACENCY];
    int overflow = 0;
    /* _init_vertex_adjacency, triangular_mesh.pyx:311-367 */
    for (i = 0; i < nv * nn; i++) adj[i] = -1;
    for (k = 0; k < nt; k++)
        for (j = 0; j < 3; j++) {
            i = tri[3 * k + j];
            if (dyn[i] < nn) adj[i * nn + dyn[i]] = k;
            dyn[i]++;
            if (dyn[i] > nn) overflow = 1;
        }
    free(dyn);
    if (overflow) { free(adj); return REF_ERR_ADJ_OVERFLOW; }
    /* _init_triangles_edges_connectivity, triangular_mesh.pyx:427-488 */
    for (i = 0; i < nt * 3; i++) triedge[i] = -1;
    for (i = 0; i < ne * 2; i++) edgetri[i] = -1;
    for (i = 0; i < ne; i++) {
        int i0 = edges[2 * i], i1 = edges[2 * i + 1];
        for (j = 0; j < nn; j++) {
            trilist[j] = adj[i0 * nn + j];
            trilist[nn + j] = adj[i1 * nn + j];
        }
        for (m = 0; m < 2 * nn - 1; m++)
            for (n = 0; n < 2 * nn - 1 - m; n++)
                if (trilist[n] > trilist[n + 1]) {
                    int t = trilist[n];
                    trilist[n] = trilist[n + 1];
                    trilist[n + 1] = t;
                }
        j = 0;
        while (j < 2 * nn && trilist[j] == -1) j++;
        while (j < 2 * nn - 1) {
            if (trilist[j] != -1 && trilist[j] == trilist[j + 1]) {
                k = trilist[j];
                if (edgetri[2 * i] == -1) edgetri[2 * i] = k;
                else edgetri[2 * i + 1] = k;
                if (triedge[3 * k] == -1) triedge[3 * k] = i;
                else if (triedge[3 * k + 1] == -1) triedge[3 * k + 1] = i;
                else triedge[3 * k + 2] = i;
                j++;
            }
            j++;
        }
    }
    free(adj);
    return REF_OK;
}

int ref_mesh_properties(int nt, const int *tri, const double *x, const double *y,
                        double *area, double *det, double *gradx, double *grady)
{   /* triangular_mesh.pyx:560-600 */
    for (int k = 0; k < nt; k++) {
        int v0 = tri[3 * k], v1 = tri[3 * k + 1], v2 = tri[3 * k + 2];
        double xa = x[v0], ya = y[v0], xb = x[v1], yb = y[v1], xc = x[v2], yc = y[v2];
        double a = 0.5 * ((xb - xa) * (yc - ya) - (yb - ya) * (xc - xa));
        area[k] = a;
        if (fabs(a) < EPSILON) return -1 - k;
        double alpha = 0.5 / a;
        gradx[3 * k] = (yb - yc) * alpha;
        gradx[3 * k + 1] = (yc - ya) * alpha;
        gradx[3 * k + 2] = (ya - yb) * alpha;
        grady[3 * k] = (xc - xb) * alpha;
        grady[3 * k + 1] = (xa - xc) * alpha;
        grady[3 * k + 2] = (xb - xa) * alpha;
        det[3 * k] = (xb * yc - xc * yb) * alpha;
        det[3 * k + 1] = (xc * ya - xa * yc) * alpha;
        det[3 * k + 2] = (xa * yb - xb * ya) * alpha;
    }
    return REF_OK;
}

void ref_edge_labels(int ne, const int *edges, const int *edgetri, const int *vlab, int *lab)
{   /* triangular_mesh.pyx:550-558 (vlab NULL: every boundary vertex is a wall) */
    for (int i = 0; i < ne; i++) {
        lab[i] = -1;
        if (edgetri[2 * i + 1] == -1) {
            int i0 = edges[2 * i], i1 = edges[2 * i + 1];
            lab[i] = BND_WALL_REF;
            if (vlab && vlab[i0] == vlab[i1]) lab[i] = vlab[i0];
        }
    }
}

/* ========================================================================== */
/* Geometry -- cylag/geom/utils.pyx, intersections.pyx, xylocalizer.pyx       */
/* ========================================================================== */
static inline int wrapk(const ref_mesh *m, int k) { return k < 0 ? k + m->nt : k; }

static inline double pintri(double x, double y, double xa, double ya, double xb, double yb,
                            double xc, double yc)
{   /* utils.pyx:332-357 */
    return ((xb - xa) * (y - ya) - (x - xa) * (yb - ya)) * ((x - xa) * (yc - ya) - (xc - xa) * (y - ya));
}

static inline int point_in_triangle_dotp(double x, double y, double x0, double y0, double x1,
                                         double y1, double x2, double y2)
{   /* utils.pyx:359-385 */
    return pintri(x, y, x0, y0, x1, y1, x2, y2) >= 0 && pintri(x, y, x1, y1, x2, y2, x0, y0) >= 0 &&
           pintri(x, y, x2, y2, x0, y0, x1, y1) >= 0;
}

int ref_point_in_tri(const ref_mesh *m, int k, double x, double y)
{   /* xylocalizer.pyx:126-167; k = -1 reads the last triangle (wraparound) */
    int kk = wrapk(m, k);
    int v0 = m->tri[3 * kk], v1 = m->tri[3 * kk + 1], v2 = m->tri[3 * kk + 2];
    if (point_in_triangle_dotp(x, y, m->x[v0], m->y[v0], m->x[v1], m->y[v1], m->x[v2], m->y[v2]))
        return k;
    return -1;
}

void ref_locate_candidates(const ref_mesh *m, int n, const double *xs, const double *ys, int k,
                           const long long *cand, int *out)
{   /* xy_localize_point_kdtree, xylocalizer.pyx:255-286: the k nearest centroids
     * (cand[i*k ...], in the KD-tree's order) are tested with point_in_tri; a
     * point none of them holds takes the expanding search */
    for (int i = 0; i < n; i++) {
        int found = -1;
        for (int j = 0; j < k && found == -1; j++) found = ref_point_in_tri(m, (int)cand[(long)i * k + j], xs[i], ys[i]);
        out[i] = found != -1 ? found : ref_localize_kdtree(m, xs[i], ys[i], k);
    }
}

static inline int orientation(const double *v0, const double *v1, const double *v2)
{   /* utils.pyx:228-252 */
    double delta = (v0[1] - v1[1]) * (v2[0] - v0[0]) - (v0[0] - v1[0]) * (v2[1] - v0[1]);
    if (delta == 0.) return 0;
    if (delta > 0.) return 1;
    return 2;
}

static inline int on_segment(const double *v0, const double *v1, const double *v2)
{   /* utils.pyx:282-303 */
    return v2[0] <= fmax(v1[0], v0[0]) && v2[0] >= fmin(v1[0], v0[0]) && v2[1] <= fmax(v1[1], v0[1]) &&
           v2[1] >= fmin(v1[1], v0[1]);
}

int ref_intersec2d(const double *p0, const double *p1, const double *s0, const double *s1)
{   /* intersections.pyx:106-148 */
    int o1 = orientation(p0, p1, s0), o2 = orientation(p0, p1, s1);
    int o3 = orientation(s0, s1, p0), o4 = orientation(s0, s1, p1);
    if (o1 != o2 && o3 != o4) return 1;
    if (o1 == 0 && on_segment(p0, p1, s0)) return 1;
    if (o2 == 0 && on_segment(p0, p1, s1)) return 1;
    if (o3 == 0 && on_segment(s0, s1, p0)) return 1;
    if (o4 == 0 && on_segment(s0, s1, p1)) return 1;
    return 0;
}

int ref_localize_brute(const ref_mesh *m, double x, double y)
{   /* xylocalizer.pyx:169-198 */
    for (int k = 0; k < m->nt; k++)
        if (ref_point_in_tri(m, k, x, y) != -1) return k;
    return -1;
}

int ref_localize_kdtree(const ref_mesh *m, double x, double y, int k_nearest)
{
    /* xylocalizer.pyx:200-286.  cKDTree returns centroids by increasing
     * distance; candidates are checked in that order among the nearest
     * k_nearest, then 2x, 5x, 10x (i.e. 10, 20, 100, 1000), then brute force.
     * Equivalent restatement: the containing triangle of least (distance,
     * index) wins if its rank is below the last expansion size, else the
     * lowest-index containing triangle.  Ties in distance are broken by index
     * (cKDTree's own tie order is not specified). */
    int nt = m->nt, best = -1, k;
    double dbest = 0.;
    if (!m->centroids) return ref_localize_brute(m, x, y);
    for (k = 0; k < nt; k++) {
        if (ref_point_in_tri(m, k, x, y) == -1) continue;
        double dx = x - m->centroids[2 * k], dy = y - m->centroids[2 * k + 1];
        double d = dx * dx + dy * dy;
        if (best == -1 || d < dbest) { best = k; dbest = d; }
    }
    if (best == -1) return -1;
    long kmax = k_nearest < nt ? k_nearest : nt;
    if (kmax < nt) {
        static const int mult[3] = {2, 5, 10};
        for (int t = 0; t < 3; t++) {
            long ke = kmax * mult[t] < nt ? kmax * mult[t] : nt;
            if (ke > kmax) kmax = ke;
        }
    }
    long rank = 0;
    for (k = 0; k < nt; k++) {
        double dx = x - m->centroids[2 * k], dy = y - m->centroids[2 * k + 1];
        double d = dx * dx + dy * dy;
        if (d < dbest || (d == dbest && k < best)) rank++;
    }
    if (rank < kmax) return best;
    return ref_localize_brute(m, x, y);
}

static int localize_from_path(const ref_mesh *m, int k, double p0x, double p0y, double x,
                              double y, int *bnd_j)
{   /* xylocalizer.pyx:338-435 */
    int new_k = -1, niter = 0, niter_max = 50, inter_edge = -1;
    double P0[2] = {p0x, p0y}, P1[2] = {x, y}, E0[2], E1[2];
    *bnd_j = -1;
    while (new_k == -1 && niter <= niter_max) {
        int intersects = 0;
        int kk = wrapk(m, k);
        for (int i = 0; i < 3; i++) {
            int j = m->triedge[3 * kk + i];
            if (j == inter_edge) continue;
            int i0 = m->edges[2 * j], i1 = m->edges[2 * j + 1];
            E0[0] = m->x[i0]; E0[1] = m->y[i0];
            E1[0] = m->x[i1]; E1[1] = m->y[i1];
            if (ref_intersec2d(P0, P1, E0, E1)) {
                intersects = 1;
                int k0 = m->edgetri[2 * j], k1 = m->edgetri[2 * j + 1];
                int kn = (k0 == k) ? k1 : k0;
                if (kn == -1) { *bnd_j = j; return -1; }
                new_k = ref_point_in_tri(m, kn, x, y);
                if (new_k != -1) return new_k;
                k = kn;
                inter_edge = j;
                break;
            }
        }
        if (!intersects) {
            new_k = ref_point_in_tri(m, k, x, y);
            if (new_k == -1) return -1;
            return k;
        }
        niter++;
    }
    return new_k;
}

int ref_xy_localize(const ref_mesh *m, int k, double p0x, double p0y, double x, double y,
                    int method, int *bnd_j)
{   /* xylocalizer.pyx:27-124 */
    int new_k;
    *bnd_j = -1;
    new_k = ref_point_in_tri(m, k, x, y);
    if (new_k != -1) return new_k;
    if (method == 1 && k == -1) method = 0;
    if (method == 2) {
        if (k == -1 && x == p0x && y == p0y) method = 0;
        if (k != -1 && x == p0x && y == p0y) method = 1;
    }
    if (method == 0) return ref_localize_kdtree(m, x, y, 10);
    if (method == 1) return REF_ERR_METHOD1 - 1000;   /* triangle_adjacency is never built */
    if (method == 2) return localize_from_path(m, k, p0x, p0y, x, y, bnd_j);
    new_k = localize_from_path(m, k, p0x, p0y, x, y, bnd_j);
    if (new_k == -1 && *bnd_j == -1) new_k = ref_localize_kdtree(m, x, y, 10);
    return new_k;
}

int ref_lower_layer_index(int nl, double zs, double zb, double z)
{   /* zlocalizer.pyx:51-71 ; Cython lowers double // to floor(a/b) */
    double hl, fi;
    if (nl == 1) return 0;
    hl = (zs - zb) / (nl - 1);
    if (!(hl > 0.)) return 0;
    fi = floor((z - zb) / hl);
    if (!(fi >= 0.)) return 0;
    if (!(fi <= (double)(nl - 2))) return nl - 2;
    return (int)fi;
}

int ref_lower_index(const double *arr, int n, double value, int last_idx)
{   /* zlocalizer.pyx:125-151 */
    for (int i = last_idx; i < n - 1; i++)
        if (value >= arr[i] && value < arr[i + 1]) return i;
    if (n > 1 && value >= arr[n - 1]) return n - 2;
    return 0;
}

/* ========================================================================== */
/* Field set -- cylag/core/eulerian_field_set.pyx                             */
/* ========================================================================== */
static const double *field_frame(const ref_fset *f, int fid, int rec)
{   /* _get_field_from_id + _get_field_slice(k, -1), :623-706 */
    size_t size = (size_t)f->npoints * f->nlayers;
    const double *base;
    if (fid == 0) base = f->elev;
    else if (fid == 1) base = f->u;
    else if (fid == 2) base = f->v;
    else if (fid == 3) base = f->w;
    else base = f->opt + (size_t)(fid - 4) * f->ntimes * size;
    return base + (size_t)rec * size;
}

static double *tmp_of(const ref_fset *f, int fid)
{   /* _get_tmp_from_id, :708-741 */
    size_t size = (size_t)f->npoints * f->nlayers;
    if (fid == 0) return f->tmp_elev;
    if (fid == 1) return f->tmp_u;
    if (fid == 2) return f->tmp_v;
    if (fid == 3) return f->tmp_w;
    return f->tmp_opt + (size_t)(fid - 4) * size;
}

static double *tmp_layer(const ref_fset *f, int fid, int l)
{   /* _get_tmp / _get_tmp_slice, :649-671,823-843 */
    double *t = tmp_of(f, fid);
    if (f->dim == 2 || l == -1) return t;
    return t + (size_t)l * f->npoints;
}

void ref_init_tmp(ref_fset *f)
{   /* _init_tmp, :743-776 */
    size_t size = (size_t)f->npoints * f->nlayers, i;
    for (int fid = 0; fid < f->dim + 1; fid++) {
        const double *s = field_frame(f, fid, 0);
        double *d = tmp_of(f, fid);
        for (i = 0; i < size; i++) d[i] = s[i];
    }
    for (int fid = 4; fid < 4 + f->nopt; fid++) {
        const double *s = field_frame(f, fid, 0);
        double *d = tmp_of(f, fid);
        for (i = 0; i < size; i++) d[i] = s[i];
    }
    memcpy(f->tmp_z0, f->tmp_elev, size * sizeof(double));
    memcpy(f->tmp_u0, f->tmp_u, size * sizeof(double));
    memcpy(f->tmp_v0, f->tmp_v, size * sizeof(double));
    if (f->tmp_w && f->tmp_w0) memcpy(f->tmp_w0, f->tmp_w, size * sizeof(double));
    f->c_valid = 0;
}

int ref_time_interpolation(ref_fset *f, double time)
{   /* _time_interpolation, :866-938 */
    size_t size = (size_t)f->npoints * f->nlayers, i;
    int ntmp = f->dim == 2 ? 3 : 4, rec;
    double dt, tf;
    if (f->frozen) return REF_OK;
    memcpy(f->tmp_z0, f->tmp_elev, size * sizeof(double));
    memcpy(f->tmp_u0, f->tmp_u, size * sizeof(double));
    memcpy(f->tmp_v0, f->tmp_v, size * sizeof(double));
    if (f->tmp_w && f->tmp_w0) memcpy(f->tmp_w0, f->tmp_w, size * sizeof(double));
    f->prev_valid = f->interp_done;
    f->interp_done = 1;
    f->record = ref_lower_index(f->times, f->ntimes, time, f->record);
    rec = f->record;
    dt = f->times[rec] - f->times[rec + 1];
    if (fabs(dt) < EPSILON) return REF_ERR_BAD_SCHEME;
    tf = (time - f->times[rec + 1]) / dt;
    for (int fid = 0; fid < ntmp; fid++) {
        const double *f0 = field_frame(f, fid, rec), *f1 = field_frame(f, fid, rec + 1);
        double *t = tmp_of(f, fid);
        for (i = 0; i < size; i++) t[i] = tf * f0[i] + (1. - tf) * f1[i];
    }
    for (int fid = 4; fid < 4 + f->nopt; fid++) {
        const double *f0 = field_frame(f, fid, rec), *f1 = field_frame(f, fid, rec + 1);
        double *t = tmp_of(f, fid);
        for (i = 0; i < size; i++) t[i] = tf * f0[i] + (1. - tf) * f1[i];
    }
    return REF_OK;
}

static inline void p1_factors(const ref_mesh *m, int k, double x, double y, double *f0,
                              double *f1, double *f2)
{
    const double *d = m->det + 3 * k, *gx = m->gradx + 3 * k, *gy = m->grady + 3 * k;
    *f0 = d[0] + x * gx[0] + y * gy[0];
    *f1 = d[1] + x * gx[1] + y * gy[1];
    *f2 = d[2] + x * gx[2] + y * gy[2];
}

static void cached_factors(ref_fset *f, const ref_mesh *m, int k, double x, double y, double *f0,
                           double *f1, double *f2)
{   /* _use_cached_factors / _cache_interp_factors, :778-821 */
    if (f->c_valid && f->c_tri == k && fabs(f->c_x - x) < EPSILON && fabs(f->c_y - y) < EPSILON) {
        *f0 = f->c_f0; *f1 = f->c_f1; *f2 = f->c_f2;
        return;
    }
    p1_factors(m, k, x, y, f0, f1, f2);
    f->c_tri = k; f->c_x = x; f->c_y = y;
    f->c_f0 = *f0; f->c_f1 = *f1; f->c_f2 = *f2;
    f->c_valid = 1;
}

int ref_interp_field_2d(ref_fset *f, const ref_mesh *m, int fid, int k, int l, double x, double y,
                        double *out)
{   /* _interpolate_field_2d, :994-1055 */
    double f0, f1, f2;
    const double *fl = tmp_layer(f, fid, f->dim == 2 ? -1 : l);
    if (k == -1) return REF_ERR_NOT_IN_TRI;
    cached_factors(f, m, k, x, y, &f0, &f1, &f2);
    const int *t = m->tri + 3 * k;
    *out = fl[t[0]] * f0 + fl[t[1]] * f1 + fl[t[2]] * f2;
    return REF_OK;
}

int ref_interp_velocity_2d(ref_fset *f, const ref_mesh *m, int k, double x, double y, double *ux,
                           double *uy)
{   /* _interpolate_velocity_2d, :1057-1115 */
    double f0, f1, f2;
    if (k == -1) return REF_ERR_NOT_IN_TRI;
    const double *u = tmp_of(f, f->fid_ux), *v = tmp_of(f, f->fid_uy);
    cached_factors(f, m, k, x, y, &f0, &f1, &f2);
    const int *t = m->tri + 3 * k;
    *ux = u[t[0]] * f0 + u[t[1]] * f1 + u[t[2]] * f2;
    *uy = v[t[0]] * f0 + v[t[1]] * f1 + v[t[2]] * f2;
    return REF_OK;
}

int ref_z_localize(ref_fset *f, const ref_mesh *m, int k, double x, double y, double z)
{   /* z_localize, :1117-1182 */
    double f0, f1, f2, zbp, zsp;
    const double *zb, *zs;
    int ll;
    if (k == -1) return PART_LOC_O;
    const int *t = m->tri + 3 * k;
    p1_factors(m, k, x, y, &f0, &f1, &f2);
    if (f->dim == 2) {
        zb = tmp_layer(f, f->fid_zb, -1);
        zs = tmp_layer(f, f->fid_zs, -1);
    } else {
        zb = tmp_layer(f, f->fid_zs, 0);
        zs = tmp_layer(f, f->fid_zs, f->nlayers - 1);
    }
    zbp = zb[t[0]] * f0 + zb[t[1]] * f1 + zb[t[2]] * f2;
    zsp = zs[t[0]] * f0 + zs[t[1]] * f1 + zs[t[2]] * f2;
    ll = f->dim == 2 ? 0 : ref_lower_layer_index(f->nlayers, zsp, zbp, z);
    if (z < zbp) ll = PART_LOC_B;
    else if (z > zsp) ll = PART_LOC_A;
    return ll;
}

int ref_interp_field_3d(ref_fset *f, const ref_mesh *m, int fid, int k, int ll, int zloc,
                        double x, double y, double z, double *out)
{   /* _interpolate_field_3d, :1184-1287 */
    int nl = f->nlayers, np = f->npoints;
    double f0, f1, f2;
    if (k == -1) return REF_ERR_NOT_IN_TRI;
    const int *t = m->tri + 3 * k;
    int i0 = t[0], i1 = t[1], i2 = t[2];
    p1_factors(m, k, x, y, &f0, &f1, &f2);
    if (zloc) {
        const double *zb = tmp_layer(f, f->fid_zs, 0), *zs = tmp_layer(f, f->fid_zs, nl - 1);
        double zbp = zb[i0] * f0 + zb[i1] * f1 + zb[i2] * f2;
        double zsp = zs[i0] * f0 + zs[i1] * f1 + zs[i2] * f2;
        ll = ref_lower_layer_index(nl, zsp, zbp, z);
    }
    if (ll == PART_LOC_B) return REF_ERR_BELOW_BOTTOM;
    if (ll == PART_LOC_A) return REF_ERR_ABOVE_SURFACE;
    const double *el = tmp_of(f, f->fid_zs), *fd = tmp_of(f, fid);
    int jl = ll * np, ju = (ll + 1) * np;
    double zlp = el[jl + i0] * f0 + el[jl + i1] * f1 + el[jl + i2] * f2;
    double zup = el[ju + i0] * f0 + el[ju + i1] * f1 + el[ju + i2] * f2;
    double dz = zlp - zup;
    if (fabs(dz) < EPSILON) return REF_ERR_DEGENERATE_LAYER;
    double zf0 = (z - zup) / dz, zf1 = (zlp - z) / dz;
    double resl = fd[jl + i0] * f0 + fd[jl + i1] * f1 + fd[jl + i2] * f2;
    double resu = fd[ju + i0] * f0 + fd[ju + i1] * f1 + fd[ju + i2] * f2;
    *out = zf0 * resl + zf1 * resu;
    return REF_OK;
}

int ref_interp_velocity_3d(ref_fset *f, const ref_mesh *m, int k, int ll, int zloc, double x,
                           double y, double z, double *ux, double *uy, double *uz)
{   /* _interpolate_velocity_3d, :1289-1394 */
    int nl = f->nlayers, np = f->npoints;
    double f0, f1, f2;
    if (k == -1) return REF_ERR_NOT_IN_TRI;
    const int *t = m->tri + 3 * k;
    int i0 = t[0], i1 = t[1], i2 = t[2];
    p1_factors(m, k, x, y, &f0, &f1, &f2);
    if (zloc) {
        const double *zb = tmp_layer(f, f->fid_zs, 0), *zs = tmp_layer(f, f->fid_zs, nl - 1);
        double zbp = zb[i0] * f0 + zb[i1] * f1 + zb[i2] * f2;
        double zsp = zs[i0] * f0 + zs[i1] * f1 + zs[i2] * f2;
        ll = ref_lower_layer_index(nl, zsp, zbp, z);
    }
    if (ll == PART_LOC_B) return REF_ERR_BELOW_BOTTOM;
    if (ll == PART_LOC_A) return REF_ERR_ABOVE_SURFACE;
    const double *el = tmp_of(f, f->fid_zs), *u = tmp_of(f, f->fid_ux), *v = tmp_of(f, f->fid_uy),
                 *w = tmp_of(f, f->fid_uz);
    int jl = ll * np, ju = (ll + 1) * np;
    double zlp = el[jl + i0] * f0 + el[jl + i1] * f1 + el[jl + i2] * f2;
    double zup = el[ju + i0] * f0 + el[ju + i1] * f1 + el[ju + i2] * f2;
    double dz = zlp - zup;
    if (fabs(dz) < EPSILON) return REF_ERR_DEGENERATE_LAYER;
    double zf0 = (z - zup) / dz, zf1 = (zlp - z) / dz;
    *ux = zf0 * (u[jl + i0] * f0 + u[jl + i1] * f1 + u[jl + i2] * f2) +
          zf1 * (u[ju + i0] * f0 + u[ju + i1] * f1 + u[ju + i2] * f2);
    *uy = zf0 * (v[jl + i0] * f0 + v[jl + i1] * f1 + v[jl + i2] * f2) +
          zf1 * (v[ju + i0] * f0 + v[ju + i1] * f1 + v[ju + i2] * f2);
    *uz = zf0 * (w[jl + i0] * f0 + w[jl + i1] * f1 + w[jl + i2] * f2) +
          zf1 * (w[ju + i0] * f0 + w[ju + i1] * f1 + w[ju + i2] * f2);
    return REF_OK;
}

static inline double k_eps_diff(double k, double eps, double fac)
{   /* _k_eps_diffusivity, :1995-1997 */
    return fac * k * k / fmax(eps, EPSILON);
}

int ref_diffusivity_gradient(ref_fset *f, const ref_mesh *m, int k, int ll, double x, double y,
                             double z, double sigc, double *Kp, double *dkdx, double *dkdy,
                             double *dkdz)
{   /* _interpolate_diffusivity_gradient_c, :1644-1763 */
    double fac = CMU / sigc, f0, f1, f2, kn0, kn1, kn2;
    if (k == -1) return REF_ERR_NOT_IN_TRI;
    if (f->fid_turbeng == -1 || f->fid_dissip == -1) return REF_ERR_NOT_IN_TRI;
    const int *t = m->tri + 3 * k;
    const double *gx = m->gradx + 3 * k, *gy = m->grady + 3 * k;
    int i0 = t[0], i1 = t[1], i2 = t[2];
    p1_factors(m, k, x, y, &f0, &f1, &f2);
    const double *kt = tmp_of(f, f->fid_turbeng), *ep = tmp_of(f, f->fid_dissip);
    if (f->dim == 2) {
        kn0 = k_eps_diff(kt[i0], ep[i0], fac);
        kn1 = k_eps_diff(kt[i1], ep[i1], fac);
        kn2 = k_eps_diff(kt[i2], ep[i2], fac);
        double K = kn0 * f0 + kn1 * f1 + kn2 * f2;
        *dkdx = kn0 * gx[0] + kn1 * gx[1] + kn2 * gx[2];
        *dkdy = kn0 * gy[0] + kn1 * gy[1] + kn2 * gy[2];
        *Kp = fmax(0., K);
        *dkdz = 0.;
        return REF_OK;
    }
    if (ll == PART_LOC_B) return REF_ERR_BELOW_BOTTOM;
    if (ll == PART_LOC_A) return REF_ERR_ABOVE_SURFACE;
    int np = f->npoints, jl = ll * np, ju = (ll + 1) * np;
    const double *el = tmp_of(f, f->fid_zs);
    kn0 = k_eps_diff(kt[jl + i0], ep[jl + i0], fac);
    kn1 = k_eps_diff(kt[jl + i1], ep[jl + i1], fac);
    kn2 = k_eps_diff(kt[jl + i2], ep[jl + i2], fac);
    double Kl = kn0 * f0 + kn1 * f1 + kn2 * f2;
    double gxl = kn0 * gx[0] + kn1 * gx[1] + kn2 * gx[2];
    double gyl = kn0 * gy[0] + kn1 * gy[1] + kn2 * gy[2];
    double zlp = el[jl + i0] * f0 + el[jl + i1] * f1 + el[jl + i2] * f2;
    double zlx = el[jl + i0] * gx[0] + el[jl + i1] * gx[1] + el[jl + i2] * gx[2];
    double zly = el[jl + i0] * gy[0] + el[jl + i1] * gy[1] + el[jl + i2] * gy[2];
    kn0 = k_eps_diff(kt[ju + i0], ep[ju + i0], fac);
    kn1 = k_eps_diff(kt[ju + i1], ep[ju + i1], fac);
    kn2 = k_eps_diff(kt[ju + i2], ep[ju + i2], fac);
    double Ku = kn0 * f0 + kn1 * f1 + kn2 * f2;
    double gxu = kn0 * gx[0] + kn1 * gx[1] + kn2 * gx[2];
    double gyu = kn0 * gy[0] + kn1 * gy[1] + kn2 * gy[2];
    double zup = el[ju + i0] * f0 + el[ju + i1] * f1 + el[ju + i2] * f2;
    double zux = el[ju + i0] * gx[0] + el[ju + i1] * gx[1] + el[ju + i2] * gx[2];
    double zuy = el[ju + i0] * gy[0] + el[ju + i1] * gy[1] + el[ju + i2] * gy[2];
    double dz = zup - zlp;
    if (fabs(dz) < EPSILON) return REF_ERR_DEGENERATE_LAYER;
    double s = (z - zlp) / dz;
    double K = (1. - s) * Kl + s * Ku;
    *dkdz = (Ku - Kl) / dz;
    *dkdx = (1. - s) * gxl + s * gxu - *dkdz * ((1. - s) * zlx + s * zux);
    *dkdy = (1. - s) * gyl + s * gyu - *dkdz * ((1. - s) * zly + s * zuy);
    *Kp = fmax(0., K);
    return REF_OK;
}

int ref_interp_acceleration_2d(ref_fset *f, const ref_mesh *m, int k, double x, double y,
                               double dt, double ua, double va, double *ax, double *ay)
{   /* _interpolate_acceleration_2d, :1396-1464 */
    if (k == -1) return REF_ERR_NOT_IN_TRI;
    const int *t = m->tri + 3 * k;
    const double *gx = m->gradx + 3 * k, *gy = m->grady + 3 * k;
    const double *u = f->tmp_u, *v = f->tmp_v, *u0 = f->tmp_u0, *v0 = f->tmp_v0;
    int i0 = t[0], i1 = t[1], i2 = t[2];
    double dudx = u[i0] * gx[0] + u[i1] * gx[1] + u[i2] * gx[2];
    double dudy = u[i0] * gy[0] + u[i1] * gy[1] + u[i2] * gy[2];
    double dvdx = v[i0] * gx[0] + v[i1] * gx[1] + v[i2] * gx[2];
    double dvdy = v[i0] * gy[0] + v[i1] * gy[1] + v[i2] * gy[2];
    double dudt = 0., dvdt = 0., f0, f1, f2;
    if (!f->frozen && f->prev_valid) {
        p1_factors(m, k, x, y, &f0, &f1, &f2);
        dudt = ((u[i0] - u0[i0]) * f0 + (u[i1] - u0[i1]) * f1 + (u[i2] - u0[i2]) * f2) / dt;
        dvdt = ((v[i0] - v0[i0]) * f0 + (v[i1] - v0[i1]) * f1 + (v[i2] - v0[i2]) * f2) / dt;
    }
    *ax = dudt + ua * dudx + va * dudy;
    *ay = dvdt + ua * dvdx + va * dvdy;
    return REF_OK;
}

static void prism_value_gradient(const double *fld, int jl, int ju, int i0, int i1, int i2,
                                 const double *gx, const double *gy, double f0, double f1,
                                 double f2, double s, double dz, double zlx, double zly, double zux,
                                 double zuy, double *val, double *gxo, double *gyo, double *gzo)
{   /* _prism_value_gradient, :1999-2023 */
    double fl = fld[jl + i0] * f0 + fld[jl + i1] * f1 + fld[jl + i2] * f2;
    double fu = fld[ju + i0] * f0 + fld[ju + i1] * f1 + fld[ju + i2] * f2;
    double gxl = fld[jl + i0] * gx[0] + fld[jl + i1] * gx[1] + fld[jl + i2] * gx[2];
    double gyl = fld[jl + i0] * gy[0] + fld[jl + i1] * gy[1] + fld[jl + i2] * gy[2];
    double gxu = fld[ju + i0] * gx[0] + fld[ju + i1] * gx[1] + fld[ju + i2] * gx[2];
    double gyu = fld[ju + i0] * gy[0] + fld[ju + i1] * gy[1] + fld[ju + i2] * gy[2];
    double dfdz = (fu - fl) / dz;
    *val = (1. - s) * fl + s * fu;
    *gxo = (1. - s) * gxl + s * gxu - dfdz * ((1. - s) * zlx + s * zux);
    *gyo = (1. - s) * gyl + s * gyu - dfdz * ((1. - s) * zly + s * zuy);
    *gzo = dfdz;
}

int ref_interp_acceleration_3d(ref_fset *f, const ref_mesh *m, int k, int ll, double x, double y,
                               double z, double dt, double ua, double va, double wa, double *ax,
                               double *ay, double *az)
{   /* _interpolate_acceleration_3d, :1466-1579 */
    if (k == -1) return REF_ERR_NOT_IN_TRI;
    const int *t = m->tri + 3 * k;
    const double *gx = m->gradx + 3 * k, *gy = m->grady + 3 * k;
    int i0 = t[0], i1 = t[1], i2 = t[2], np = f->npoints;
    double f0, f1, f2;
    p1_factors(m, k, x, y, &f0, &f1, &f2);
    const double *el = tmp_of(f, f->fid_zs);
    int jl = ll * np, ju = (ll + 1) * np;
    double zlp = el[jl + i0] * f0 + el[jl + i1] * f1 + el[jl + i2] * f2;
    double zup = el[ju + i0] * f0 + el[ju + i1] * f1 + el[ju + i2] * f2;
    double dz = zup - zlp;
    if (fabs(dz) < EPSILON) return REF_ERR_DEGENERATE_LAYER;
    double s = (z - zlp) / dz;
    double zlx = el[jl + i0] * gx[0] + el[jl + i1] * gx[1] + el[jl + i2] * gx[2];
    double zly = el[jl + i0] * gy[0] + el[jl + i1] * gy[1] + el[jl + i2] * gy[2];
    double zux = el[ju + i0] * gx[0] + el[ju + i1] * gx[1] + el[ju + i2] * gx[2];
    double zuy = el[ju + i0] * gy[0] + el[ju + i1] * gy[1] + el[ju + i2] * gy[2];
    double un, dudx, dudy, dudz, vn, dvdx, dvdy, dvdz, wn, dwdx, dwdy, dwdz, tmp, d1, d2, d3;
    prism_value_gradient(f->tmp_u, jl, ju, i0, i1, i2, gx, gy, f0, f1, f2, s, dz, zlx, zly, zux,
                         zuy, &un, &dudx, &dudy, &dudz);
    prism_value_gradient(f->tmp_v, jl, ju, i0, i1, i2, gx, gy, f0, f1, f2, s, dz, zlx, zly, zux,
                         zuy, &vn, &dvdx, &dvdy, &dvdz);
    prism_value_gradient(f->tmp_w, jl, ju, i0, i1, i2, gx, gy, f0, f1, f2, s, dz, zlx, zly, zux,
                         zuy, &wn, &dwdx, &dwdy, &dwdz);
    double dudt = 0., dvdt = 0., dwdt = 0.;
    if (!f->frozen && f->prev_valid) {
        prism_value_gradient(f->tmp_u0, jl, ju, i0, i1, i2, gx, gy, f0, f1, f2, s, dz, zlx, zly,
                             zux, zuy, &tmp, &d1, &d2, &d3);
        dudt = (un - tmp) / dt;
        prism_value_gradient(f->tmp_v0, jl, ju, i0, i1, i2, gx, gy, f0, f1, f2, s, dz, zlx, zly,
                             zux, zuy, &tmp, &d1, &d2, &d3);
        dvdt = (vn - tmp) / dt;
        prism_value_gradient(f->tmp_w0, jl, ju, i0, i1, i2, gx, gy, f0, f1, f2, s, dz, zlx, zly,
                             zux, zuy, &tmp, &d1, &d2, &d3);
        dwdt = (wn - tmp) / dt;
    }
    *ax = dudt + ua * dudx + va * dudy + wa * dudz;
    *ay = dvdt + ua * dvdx + va * dvdy + wa * dvdz;
    *az = dwdt + ua * dwdx + va * dwdy + wa * dwdz;
    return REF_OK;
}

static double friction_calc(const ref_fset *f, double h)
{   /* friction_calc, :1957-1993 */
    double c = f->friction_coefficient, Cf, aux;
    switch (f->friction_model) {
    case 0: Cf = 2. * GRAV / pow(c, 2.0); break;
    case 1: Cf = 2. * GRAV / pow(c, 2.0) / pow(fmax(EPSILON, h), 1. / 3.); break;
    case 2: Cf = 2. * GRAV * pow(c, 2.0) / pow(fmax(EPSILON, h), 1. / 3.); break;
    case 3:
        aux = h * 11.036 / c;
        aux = aux > 1.001 ? aux : 1.001;
        Cf = 2. / pow(log(aux) / KAPPA, 2.0);
        break;
    default: Cf = NAN;
    }
    return Cf;
}

static double surface_velocity(ref_fset *f, const ref_mesh *m, int k, double x, double y,
                               double dt, double zsp)
{   /* _interpolate_surface_velocity, :1581-1642 */
    double f0, f1, f2;
    if (f->frozen || !f->prev_valid) return 0.;
    p1_factors(m, k, x, y, &f0, &f1, &f2);
    const int *t = m->tri + 3 * k;
    double z0p = f->tmp_z0[t[0]] * f0 + f->tmp_z0[t[1]] * f1 + f->tmp_z0[t[2]] * f2;
    return (zsp - z0p) / dt;
}

static int pseudo3d_corrections(ref_fset *f, const ref_mesh *m, int k, double x, double y,
                                double z, double dt, double ux, double uy, double *wout,
                                double *hor_out)
{   /* pseudo_3d_velocity_corrections, :1887-1955 */
    const double *zs_n = tmp_layer(f, f->fid_zs, -1), *zb_n = tmp_layer(f, f->fid_zb, -1);
    const double *gx = m->gradx + 3 * k, *gy = m->grady + 3 * k;
    int kk = wrapk(m, k);
    const int *t = m->tri + 3 * kk;
    int i0 = t[0], i1 = t[1], i2 = t[2], e;
    double zs, zb, h, sigma, Cf, z0, hor, uc, vc;
    e = ref_interp_field_2d(f, m, f->fid_zs, k, -1, x, y, &zs);
    if (e) return e;
    e = ref_interp_field_2d(f, m, f->fid_zb, k, -1, x, y, &zb);
    if (e) return e;
    h = fmax(EPSILON, zs - zb);
    sigma = fmin(1., fmax(0., (z - zb) / h));
    Cf = friction_calc(f, h);
    z0 = fmax(EPSILON, h * exp(-(1. + KAPPA / sqrt(Cf / 2.))));
    if (z - zb <= z0) hor = 0.0;
    else hor = log((z - zb) / z0) / (log(h / z0) - 1.);
    uc = ux * hor;
    vc = uy * hor;
    double dzsdx = zs_n[i0] * gx[0] + zs_n[i1] * gx[1] + zs_n[i2] * gx[2];
    double dzsdy = zs_n[i0] * gy[0] + zs_n[i1] * gy[1] + zs_n[i2] * gy[2];
    double dzbdx = zb_n[i0] * gx[0] + zb_n[i1] * gx[1] + zb_n[i2] * gx[2];
    double dzbdy = zb_n[i0] * gy[0] + zb_n[i1] * gy[1] + zb_n[i2] * gy[2];
    double dzsdt = surface_velocity(f, m, k, x, y, dt, zs);
    *wout = (1. - sigma) * (uc * dzbdx + vc * dzbdy) + sigma * (dzsdt + uc * dzsdx + vc * dzsdy);
    *hor_out = hor;
    return REF_OK;
}

/* ========================================================================== */
/* Particle set -- cylag/core/lagrangian_particle_set.pyx                     */
/* ========================================================================== */
void ref_delete_single(ref_pset *p, int i)
{   /* delete_single, :754-788 */
    if (p->inactive[i] != 0) return;
    int d = p->dim;
    p->inactive[i] = 1;
    p->diameter[i] = 0.;
    p->volume[i] = 0.;
    for (int c = 0; c < d; c++) {
        p->position[i * d + c] = 0.;
        p->velocity[i * d + c] = 0.;
        p->us[i * d + c] = 0.;
        p->last_position[i * d + c] = 0.;
        p->last_velocity[i * d + c] = 0.;
    }
    if (p->compute_depth && p->depth) p->depth[i] = 0.;
    p->tri[i] = -1;
    p->lowerlayer[i] = 0;
    p->last_tri[i] = -1;
    p->npart_active -= 1;
    p->tag[i] = p->maxtag + 1;
    p->maxtag += 1;
}

double ref_drag_coefficient(const ref_pset *p, double U, double dp)
{   /* drag_coefficient, :790-832 */
    double Re = fmax(EPSILON, U * dp / NU0), Cd;
    if (p->drag_coef_model == 0) Cd = p->drag_coef;
    else if (p->drag_coef_model == 1) {
        if (Re <= 1000.) Cd = (24. / Re) * (1. + 0.15 * pow(Re, 0.687));
        else Cd = 0.44;
    } else if (p->drag_coef_model == 2) {
        double phi1 = pow(24. / Re, 10.) + pow(21. * pow(Re, -0.67), 10.) +
                      pow(4. * pow(Re, -0.33), 10.) + pow(0.4, 10);
        double phi2 = 1. / (pow(0.148 * pow(Re, 0.11), -10.) + pow(0.5, -10.));
        double phi3 = pow(1.57 * pow(Re, -1.625) * 1.e8, 10.);
        double phi4 = 1. / (pow(6e-17 * pow(Re, 2.63), -10.) + pow(0.2, -10.));
        Cd = pow(1. / (pow(phi1 + phi2, -1.) + pow(phi3, -1.)) + phi4, 1 / 10.);
    } else Cd = 0.44;
    return Cd;
}

double ref_buoyancy_velocity(const ref_pset *p, double dp, int model_hr, double rho0)
{   /* buoyancy_velocity, :834-890 */
    double rhop = p->density, Cd = p->drag_coef, s = rhop / rho0, wb;
    if (p->buoyancy_velocity_model == 0) return p->buoyancy_velocity_value;
    if (dp < 1.e-4) wb = fabs(s - 1.) * (GRAV * pow(dp, 2.0)) / (18. * NU0);
    else if (dp > 1.e-3) {
        if (model_hr == 0) wb = sqrt(fabs(s - 1.) * (4. * dp * GRAV) / (3. * Cd));
        else wb = 1.1 * sqrt((s - 1.) * GRAV * dp);
    } else
        wb = (10. * NU0 / dp) * (sqrt(1. + fabs(s - 1.) * GRAV * pow(dp, 3.0) / (100. * pow(NU0, 2.0))) - 1.);
    if (s > 1.) wb *= -1.;
    return wb;
}

void ref_initialize_model_coefficients(ref_pset *p)
{   /* _initialize_model_coefficients, :349-375 */
    double rhop = p->density, Ca = p->added_mass_coef;
    if (!p->added_mass_force) {
        p->a0c = (rhop - p->rho0) / rhop;
        p->a1c = 0.5 * (p->rho0 / rhop);
        p->a2c = 0.;
        p->a3c = 1. / rhop;
    } else {
        double aux = 1. / (rhop + Ca * p->rho0);
        p->a0c = (rhop - p->rho0) * aux;
        p->a1c = 0.5 * p->rho0 * aux;
        p->a2c = p->rho0 * (1. + Ca) * aux;
        p->a3c = aux;
    }
}

/* ========================================================================== */
/* LSM-1 -- cylag/lsm/update_state_lsm1.pyx                                   */
/* ========================================================================== */
#define CK(e) do { int _e = (e); if (_e) return _e; } while (0)

/* mean velocity at a point, with pseudo-3D corrections (shared by all schemes) */
static int mean_velocity(ref_solver *s, int k, int ll, double x, double y, double z, double *ux,
                         double *uy, double *uz)
{
    ref_fset *f = s->fset;
    const ref_mesh *m = s->mesh;
    if (f->dim == 2) {
        CK(ref_interp_velocity_2d(f, m, k, x, y, ux, uy));
        *uz = 0.;
        if (s->pset->dim == 3) {
            double uzc, ucc;
            CK(pseudo3d_corrections(f, m, k, x, y, z, s->time_step, *ux, *uy, &uzc, &ucc));
            *uz += uzc;
            *ux *= ucc;
            *uy *= ucc;
        }
    } else {
        CK(ref_interp_velocity_3d(f, m, k, ll, 0, x, y, z, ux, uy, uz));
    }
    return REF_OK;
}

int ref_mean_velocity_p3d(ref_solver *s, int k, int ll, double x, double y, double z, double *ux,
                          double *uy, double *uz)
{
    return mean_velocity(s, k, ll, x, y, z, ux, uy, uz);
}

static void euler_lsm1(ref_solver *s, int i, int *err)
{   /* euler_lsm1, update_state_lsm1.pyx:135-177 */
    ref_pset *p = s->pset;
    int d = p->dim;
    double *xp = p->position + i * d, *up = p->velocity + i * d, *us = p->us + i * d;
    double dt = s->time_step, ux, uy, uz;
    *err = mean_velocity(s, p->tri[i], d == 3 ? p->lowerlayer[i] : 0, xp[0], xp[1],
                         d == 3 ? xp[2] : 0., &ux, &uy, &uz);
    if (*err) return;
    xp[0] += ux * dt;
    xp[1] += uy * dt;
    up[0] = ux; up[1] = uy;
    us[0] = ux; us[1] = uy;
    if (d == 3) { xp[2] += uz * dt; up[2] = uz; us[2] = uz; }
}

/* rk2/rk3/rk4_lsm1, update_state_lsm1.pyx:179-656.  nst stages with the given
 * stage fractions; returns bnd_j (early exit at boundary) */
static int rk_lsm1(ref_solver *s, int i, int order, int *err)
{
    ref_pset *p = s->pset;
    ref_fset *f = s->fset;
    const ref_mesh *m = s->mesh;
    int d = p->dim;
    double *xp = p->position + i * d, *up = p->velocity + i * d, *us = p->us + i * d;
    double dt = s->time_step, ux, uy, uz, xt, yt, zt = 0.;
    double kx[4] = {0}, ky[4] = {0}, kz[4] = {0};
    /* stage position fractions: RK2 {0.5}, RK3 {1/3, 2/3}, RK4 {0.5,0.5,1} */
    int nstage = order;
    int tmp_tri, bnd = -1, tmp_ll = 0;
    *err = 0;
    *err = mean_velocity(s, p->tri[i], d == 3 ? p->lowerlayer[i] : 0, xp[0], xp[1],
                         d == 3 ? xp[2] : 0., &ux, &uy, &uz);
    if (*err) return -1;
    up[0] = ux; up[1] = uy; us[0] = ux; us[1] = uy;
    if (d == 3) { up[2] = uz; us[2] = uz; }
    for (int st = 0; st < nstage - 1 + (order == 2 ? 0 : 0); st++) {
        /* k of the stage just evaluated */
        kx[st] = ux * dt; ky[st] = uy * dt; if (d == 3) kz[st] = uz * dt;
        if (order == 2) {
            /* rk2 uses 0.5*ux*dt directly (update_state_lsm1.pyx:224-227) */
            xt = xp[0] + 0.5 * ux * dt;
            yt = xp[1] + 0.5 * uy * dt;
            if (d == 3) zt = xp[2] + 0.5 * uz * dt;
        } else if (order == 3) {
            double c = st == 0 ? (1. / 3.) : (2. / 3.);
            xt = xp[0] + c * kx[st];
            yt = xp[1] + c * ky[st];
            zt = d == 3 ? xp[2] + c * kz[st] : 0.;
        } else {
            if (st < 2) {
                xt = xp[0] + 0.5 * kx[st];
                yt = xp[1] + 0.5 * ky[st];
                zt = d == 3 ? xp[2] + 0.5 * kz[st] : 0.;
            } else {
                xt = xp[0] + kx[st];
                yt = xp[1] + ky[st];
                zt = d == 3 ? xp[2] + kz[st] : 0.;
            }
        }
        p->last_tri[i] = p->tri[i];
        tmp_tri = ref_xy_localize(m, p->last_tri[i], p->last_position[i * d],
                                  p->last_position[i * d + 1], xt, yt, 3, &bnd);
        if (d == 3) tmp_ll = ref_z_localize(f, m, tmp_tri, xt, yt, zt);
        if (bnd != -1 || (d == 3 && (tmp_ll == -2 || tmp_ll == -3))) {
            p->tri[i] = tmp_tri;
            if (d == 3) p->lowerlayer[i] = tmp_ll;
            xp[0] += up[0] * dt;
            xp[1] += up[1] * dt;
            if (d == 3) xp[2] += up[2] * dt;
            return bnd;
        }
        *err = mean_velocity(s, tmp_tri, tmp_ll, xt, yt, zt, &ux, &uy, &uz);
        if (*err) return -1;
    }
    if (order == 2) {
        xp[0] += ux * dt;
        xp[1] += uy * dt;
        if (d == 3) xp[2] += uz * dt;
    } else if (order == 3) {
        kx[2] = ux * dt; ky[2] = uy * dt; if (d == 3) kz[2] = uz * dt;
        xp[0] += (0.25 * kx[0] + 0.75 * kx[2]);
        xp[1] += (0.25 * ky[0] + 0.75 * ky[2]);
        if (d == 3) xp[2] += (0.25 * kz[0] + 0.75 * kz[2]);
    } else {
        kx[3] = ux * dt; ky[3] = uy * dt; if (d == 3) kz[3] = uz * dt;
        xp[0] += (kx[0] + 2. * kx[1] + 2. * kx[2] + kx[3]) / 6.;
        xp[1] += (ky[0] + 2. * ky[1] + 2. * ky[2] + ky[3]) / 6.;
        if (d == 3) xp[2] += (kz[0] + 2. * kz[1] + 2. * kz[2] + kz[3]) / 6.;
    }
    return -1;
}

static int diffusion_common(ref_solver *s, int i, int tri0, int ll0, double *dx, double *dy,
                            double *dz)
{   /* diffusion_lsm1 (update_state_lsm1.pyx:658-726) == diffusion_lsm2
     * (update_state_lsm2.pyx:1023-1086): displacement increments */
    ref_pset *p = s->pset;
    const ref_params *par = s->par;
    int d = p->dim, dm = par->diffusion_model;
    double dt = s->time_step, sqrtdt = sqrt(dt);
    double xix = 0., xiy = 0., xiz = 0., diffx = 0., diffy = 0., diffz = 0.;
    double driftx = 0., drifty = 0., driftz = 0., Kp, dkdx, dkdy, dkdz, z0 = 0.;
    if (dm == 2 || dm == 3) {
        if (d == 3) z0 = p->last_position[i * d + 2];
        CK(ref_diffusivity_gradient(s->fset, s->mesh, tri0, ll0, p->last_position[i * d],
                                    p->last_position[i * d + 1], z0, par->schmidt_number, &Kp,
                                    &dkdx, &dkdy, &dkdz));
        diffx = sqrt(2. * Kp) * sqrtdt;
        diffy = diffx;
        driftx = dkdx * dt;
        drifty = dkdy * dt;
        if (d == 3) {
            if (dm == 2) diffz = sqrt(2. * par->vertical_diffusivity) * sqrtdt;
            else { diffz = diffx; driftz = dkdz * dt; }
        }
    } else {
        diffx = sqrt(2. * par->horizontal_diffusivity) * sqrtdt;
        diffy = diffx;
        if (d == 3) diffz = sqrt(2. * par->vertical_diffusivity) * sqrtdt;
    }
    if (diffx != 0.) CK(ref_generate_random_pair(&s->rng, par->rng_method, &xix, &xiy));
    if (d == 3 && diffz != 0.) CK(ref_generate_random_single(&s->rng, par->rng_method, &xiz));
    *dx = driftx + diffx * xix;
    *dy = drifty + diffy * xiy;
    *dz = driftz + diffz * xiz;
    return REF_OK;
}

static int update_state_lsm1(ref_solver *s, int i, int *bnd_j)
{   /* update_state_lsm1.pyx:30-133 */
    ref_pset *p = s->pset;
    const ref_params *par = s->par;
    int d = p->dim, c, err = 0;
    int tri0 = p->tri[i], ll0 = 0;
    double dt = s->time_step;
    *bnd_j = -1;
    for (c = 0; c < d; c++) {
        p->last_position[i * d + c] = p->position[i * d + c];
        p->last_velocity[i * d + c] = p->velocity[i * d + c];
    }
    if (d == 3) ll0 = p->lowerlayer[i];
    if (par->time_scheme == 1) euler_lsm1(s, i, &err);
    else if (par->time_scheme >= 2 && par->time_scheme <= 4)
        *bnd_j = rk_lsm1(s, i, par->time_scheme, &err);
    else return REF_ERR_BAD_SCHEME;
    if (err) return err;
    if (par->diffusion_model != 0) {
        double ddx, ddy, ddz;
        CK(diffusion_common(s, i, tri0, ll0, &ddx, &ddy, &ddz));
        p->position[i * d] += ddx;
        p->position[i * d + 1] += ddy;
        if (d == 3) p->position[i * d + 2] += ddz;
    }
    if (d == 3) {
        double wb = ref_buoyancy_velocity(p, p->diameter[i], 0, p->rho0);
        p->position[i * d + 2] += wb * dt;
        p->velocity[i * d + 2] += wb;
    }
    if (*bnd_j == -1) {
        p->last_tri[i] = p->tri[i];
        p->tri[i] = ref_xy_localize(s->mesh, p->last_tri[i], p->last_position[i * d],
                                    p->last_position[i * d + 1], p->position[i * d],
                                    p->position[i * d + 1], 3, bnd_j);
        if (d == 3)
            p->lowerlayer[i] = ref_z_localize(s->fset, s->mesh, p->tri[i], p->position[i * d],
                                              p->position[i * d + 1], p->position[i * d + 2]);
    }
    return REF_OK;
}

/* ========================================================================== */
/* Model 0 -- cylag/lsm/update_state_test.pyx:28-87                           */
/* ========================================================================== */
static int update_state_test(ref_solver *s, int i, int *bnd_j)
{
    ref_pset *p = s->pset;
    int d = p->dim, c;
    double dt = s->time_step;
    for (c = 0; c < d; c++) {
        p->last_position[i * d + c] = p->position[i * d + c];
        p->last_velocity[i * d + c] = p->velocity[i * d + c];
    }
    for (c = 0; c < d; c++) p->position[i * d + c] += p->velocity[i * d + c] * dt;
    p->last_tri[i] = p->tri[i];
    p->tri[i] = ref_xy_localize(s->mesh, p->last_tri[i], p->last_position[i * d],
                                p->last_position[i * d + 1], p->position[i * d],
                                p->position[i * d + 1], 2, bnd_j);
    if (p->tri[i] < -1) return REF_ERR_METHOD1;
    if (d == 3)
        p->lowerlayer[i] = ref_z_localize(s->fset, s->mesh, p->tri[i], p->position[i * d],
                                          p->position[i * d + 1], p->position[i * d + 2]);
    return REF_OK;
}

/* LSM-2 and LSM-3 bodies live in cpu_ref_lsm23.c */
int ref_update_state_lsm2(ref_solver *s, int i, int *bnd_j);
int ref_update_state_lsm3(ref_solver *s, int i, int *bnd_j);
int ref_lsm2_fluctuations(ref_solver *s, int i, double *f1, double *f2, double *f3);
int ref_lsm3_fluctuations(ref_solver *s, int i, double *f1, double *f2, double *f3);

void ref_symmetric_point2d(const double *p0, const double *p1, const double *pt, double damp,
                           double *out)
{   /* compute_symetric_point2d, intersections.pyx:27-62 */
    double n0 = p0[1] - p1[1], n1 = p1[0] - p0[0];
    double norm2 = n0 * n0 + n1 * n1;
    double u0 = pt[0] - p0[0], u1 = pt[1] - p0[1];
    double dist = u0 * n0 + u1 * n1;
    double coef = 1. - damp / 2.;
    out[0] = pt[0] - 2. * ((dist * coef) / norm2) * n0;
    out[1] = pt[1] - 2. * ((dist * coef) / norm2) * n1;
}

/* ========================================================================== */
/* Boundary conditions -- cylag/boundary_conditions/compute_bc.pyx:35-395     */
/* ========================================================================== */
static int compute_bc(ref_solver *s, int i, int bnd_j)
{
    ref_pset *p = s->pset;
    ref_fset *f = s->fset;
    const ref_mesh *m = s->mesh;
    const ref_params *par = s->par;
    int d = p->dim, bnd_jj, rc = 0, strand_condition = 0;
    double damp = fmin(par->rebound_damping_coef, 1.0 - EPSILON_REBOUND);
    double strand = par->stranding_probability;
    double rough = par->wall_roughness_angle * DEG_TO_RAD, max_theta = 85.0 * DEG_TO_RAD;
    const int *lab = m->edge_labels;
    if (bnd_j != -1) {
        if (lab[bnd_j] == BND_WALL_REF && strand > 0.)
            strand_condition = ref_random_uniform(&s->rng) < strand;
        if (par->stranding_output && strand_condition) {
            if (p->stranded_count < p->max_stranded) {
                double zv = d == 2 ? 0.0 : p->position[i * d + 2];
                p->stranded_positions[3 * p->stranded_count] = p->position[i * d];
                p->stranded_positions[3 * p->stranded_count + 1] = p->position[i * d + 1];
                p->stranded_positions[3 * p->stranded_count + 2] = zv;
                p->stranded_count++;
            }
        }
        if (lab[bnd_j] == BND_WALL_REF && !strand_condition) {
            bnd_jj = bnd_j;
            rc = 0;
            while (p->tri[i] == PART_LOC_O && rc < par->max_wall_rebounds) {
                int i0 = m->edges[2 * bnd_jj], i1 = m->edges[2 * bnd_jj + 1];
                double v0[2] = {m->x[i0], m->y[i0]}, v1[2] = {m->x[i1], m->y[i1]};
                double sym[2];
                ref_symmetric_point2d(v0, v1, p->position + i * d, damp, sym);
                double px = sym[0], py = sym[1];
                p->position[i * d] = px;
                p->position[i * d + 1] = py;
                p->tri[i] = ref_xy_localize(m, p->last_tri[i], p->last_position[i * d],
                                            p->last_position[i * d + 1], px, py, 3, &bnd_jj);
                if (p->tri[i] == PART_LOC_O && bnd_jj != -1 && lab[bnd_jj] >= BND_OPEN_REF) {
                    if (par->bnd_statistics) s->bnd_stats[lab[bnd_jj] - 1] += 1;
                    ref_delete_single(p, i);
                    break;
                }
                if (par->model == 0 || par->model >= 2 || par->model == 3) {
                    /* normal2u, utils.pyx:99-119 */
                    double nrm = sqrt(pow(v1[0] - v0[0], 2.0) + pow(v1[1] - v0[1], 2.0));
                    double nn[2] = {(v0[1] - v1[1]) / nrm, (v1[0] - v0[0]) / nrm};
                    double nnv[2] = {nn[0], nn[1]};
                    if (par->model == 0 || par->model >= 2) {
                        if (par->boundary_conditions_type == 2 && rough > 0.) {
                            double th;
                            CK(ref_generate_random_single(&s->rng, par->rng_method, &th));
                            th *= rough;
                            th = fmax(-max_theta, fmin(max_theta, th));
                            double ct = cos(th), st = sin(th);
                            double a = nnv[0] * ct - nnv[1] * st, b = nnv[0] * st + nnv[1] * ct;
                            nnv[0] = a; nnv[1] = b;
                        }
                        double *vel = p->velocity + i * d;
                        double up0 = vel[0];
                        vel[0] = up0 * nnv[0] + vel[1] * nnv[1];
                        vel[1] = -up0 * nnv[1] + vel[1] * nnv[0];
                        vel[0] = -1. * vel[0];
                        up0 = vel[0];
                        vel[0] = up0 * nnv[0] - vel[1] * nnv[1];
                        vel[1] = up0 * nnv[1] + vel[1] * nnv[0];
                        vel[0] *= 1. - damp;
                        vel[1] *= 1. - damp;
                        if (d == 3) vel[2] *= 1. - damp;
                    }
                    if (par->model == 3) {
                        nnv[0] = nn[0]; nnv[1] = nn[1];
                        if (par->boundary_conditions_type == 2 && rough > 0.) {
                            double th;
                            CK(ref_generate_random_single(&s->rng, par->rng_method, &th));
                            th *= rough;
                            th = fmax(-max_theta, fmin(max_theta, th));
                            double ct = cos(th), st = sin(th);
                            double a = nnv[0] * ct - nnv[1] * st, b = nnv[0] * st + nnv[1] * ct;
                            nnv[0] = a; nnv[1] = b;
                        }
                        double *vs = p->us + i * d;
                        double up0 = vs[0];
                        vs[0] = up0 * nnv[0] + vs[1] * nnv[1];
                        vs[1] = -up0 * nnv[1] + vs[1] * nnv[0];
                        vs[0] = -1. * vs[0];
                        up0 = vs[0];
                        vs[0] = up0 * nnv[0] - vs[1] * nnv[1];
                        vs[1] = up0 * nnv[1] + vs[1] * nnv[0];
                    }
                }
                rc++;
                if (p->tri[i] == PART_LOC_O && bnd_jj == -1) break;
            }
            if (d == 3 && p->tri[i] != PART_LOC_O)
                p->lowerlayer[i] = ref_z_localize(f, m, p->tri[i], p->position[i * d],
                                                  p->position[i * d + 1], p->position[i * d + 2]);
        }
        if (lab[bnd_j] >= BND_OPEN_REF) {
            ref_delete_single(p, i);
            if (par->bnd_statistics) s->bnd_stats[lab[bnd_j] - 1] += 1;
        }
    }
    if (d == 3) {
        int below = p->lowerlayer[i] == PART_LOC_B && p->tri[i] != PART_LOC_O;
        int above = 0;
        for (int pass = 0; pass < 2; pass++) {
            int isb = pass == 0;
            if (isb && !below) continue;
            if (!isb) {
                above = p->lowerlayer[i] == PART_LOC_A && p->tri[i] != PART_LOC_O;
                if (!above) continue;
            }
            double zbp, zsp;
            double x = p->position[i * d], y = p->position[i * d + 1];
            if (f->dim == 2) {
                CK(ref_interp_field_2d(f, m, f->fid_zb, p->tri[i], 0, x, y, &zbp));
                CK(ref_interp_field_2d(f, m, f->fid_zs, p->tri[i], 0, x, y, &zsp));
            } else {
                CK(ref_interp_field_2d(f, m, f->fid_zs, p->tri[i], 0, x, y, &zbp));
                CK(ref_interp_field_2d(f, m, f->fid_zs, p->tri[i], f->nlayers - 1, x, y, &zsp));
            }
            double *z = p->position + i * d + 2;
            double hp = fmax(0., zsp - zbp);
            if (isb && par->bottom_deposition) {
                /* Deposition on the bed: no counterpart in the reference, which
                   reflects (the branch below, compute_bc.pyx:294-332).  The particle
                   is recorded at (x, y, bed elevation) with the iteration of the
                   step, then deleted like any particle leaving the water column. */
                if (p->deposited_count < p->max_deposited) {
                    int q = p->deposited_count++;
                    p->deposited_positions[3 * q] = x;
                    p->deposited_positions[3 * q + 1] = y;
                    p->deposited_positions[3 * q + 2] = zbp;
                    p->deposited_iteration[q] = s->iteration + 1;
                }
                ref_delete_single(p, i);
                break;
            }
            if (isb) {
                double dzz = fabs(zbp - *z);
                *z = fmin(zbp + (1. - damp) * dzz, zbp + 0.5 * hp);
                p->lowerlayer[i] = ref_lower_layer_index(f->nlayers, zsp, zbp, *z);
                if (par->model >= 2) p->velocity[i * d + 2] = fabs(p->velocity[i * d + 2]) * (1. - damp);
                if (par->model == 3) p->us[i * d + 2] = fabs(p->us[i * d + 2]) * (1. - damp);
            } else {
                double dzz = fabs(*z - zsp);
                *z = fmax(zsp - (1. - damp) * dzz, zsp - 0.5 * hp);
                p->lowerlayer[i] = ref_lower_layer_index(f->nlayers, zsp, zbp, *z);
                if (par->model >= 2) p->velocity[i * d + 2] = -fabs(p->velocity[i * d + 2]) * (1. - damp);
                if (par->model == 3) p->us[i * d + 2] = -fabs(p->us[i * d + 2]) * (1. - damp);
            }
        }
    }
    if (p->inactive[i] == 0) {
        if (p->tri[i] == PART_LOC_O) ref_delete_single(p, i);
        if (p->lowerlayer[i] == PART_LOC_B) ref_delete_single(p, i);
        if (p->lowerlayer[i] == PART_LOC_A) ref_delete_single(p, i);
    }
    return REF_OK;
}

/* ========================================================================== */
/* Solver -- cylag/core/solver.pyx                                            */
/* ========================================================================== */
int ref_initialize_particle_state(ref_solver *s, int i)
{   /* initialize_particle_state, solver.pyx:238-380 */
    ref_pset *p = s->pset;
    ref_fset *f = s->fset;
    const ref_mesh *m = s->mesh;
    const ref_params *par = s->par;
    int d = p->dim, bj;
    double *xp = p->position + i * d, *up = p->velocity + i * d, *us = p->us + i * d;
    double zs, zb, ux, uy, uz, f1, f2, f3;
    if (s->has_trifinder && s->init_tri) p->tri[i] = s->init_tri[i];
    else p->tri[i] = ref_xy_localize(m, -1, 0., 0., xp[0], xp[1], 0, &bj);
    if (p->tri[i] == -1) ref_delete_single(p, i);
    if (d == 3 && p->inactive[i] == 0) {
        CK(ref_interp_field_2d(f, m, f->fid_zs, p->tri[i], f->nlayers - 1, xp[0], xp[1], &zs));
        if (f->dim == 2) CK(ref_interp_field_2d(f, m, f->fid_zb, p->tri[i], 0, xp[0], xp[1], &zb));
        else CK(ref_interp_field_2d(f, m, f->fid_zs, p->tri[i], 0, xp[0], xp[1], &zb));
        p->lowerlayer[i] = ref_lower_layer_index(f->nlayers, zs, zb, xp[2]);
        if (xp[2] < zb) xp[2] = zb + EPSILON;
        if (xp[2] > zs) xp[2] = zs - EPSILON;
    }
    if (p->inactive[i] == 0) {
        if (p->compute_depth && p->depth) {
            CK(ref_interp_field_2d(f, m, f->fid_zs, p->tri[i], f->nlayers - 1, xp[0], xp[1], &zs));
            p->depth[i] = zs - xp[2];
        }
        if (par->particle_velocity_init == 1) {
            if (f->dim == 2) {
                CK(ref_interp_velocity_2d(f, m, p->tri[i], xp[0], xp[1], &ux, &uy));
                up[0] = ux; up[1] = uy; us[0] = ux; us[1] = uy;
            } else {
                CK(ref_interp_velocity_3d(f, m, p->tri[i], p->lowerlayer[i], 0, xp[0], xp[1],
                                          xp[2], &ux, &uy, &uz));
                up[0] = ux; up[1] = uy; up[2] = uz;
                us[0] = ux; us[1] = uy; us[2] = uz;
            }
            if (par->model == 2 && par->diffusion_lsm2_option >= 2 && par->diffusion_model != 0) {
                CK(ref_lsm2_fluctuations(s, i, &f1, &f2, &f3));
                up[0] += f1;
                up[1] += f2;
                if (d == 3) up[2] += f3;
            }
            for (int c = 0; c < d; c++) us[c] = up[c];
            if (par->model == 3) {
                CK(ref_lsm3_fluctuations(s, i, &f1, &f2, &f3));
                us[0] += f1;
                us[1] += f2;
                if (d == 3) us[2] += f3;
            }
        }
    }
    return REF_OK;
}

static int csec_cross(const ref_solver *s, int c, double x0, double y0, double x1, double y1)
{   /* ControlSection.c_cross, control_shapes.pyx:56-85 */
    const double *g = s->csec_seg;
    for (int q = 0; q < c; q++) g += 4 * s->csec_nseg[q];
    for (int k = 0; k < s->csec_nseg[c]; k++, g += 4) {
        double p0[2] = {x0, y0}, p1[2] = {x1, y1}, s0[2] = {g[0], g[1]}, s1[2] = {g[2], g[3]};
        if (ref_intersec2d(p0, p1, s0, s1)) return 1;
    }
    return 0;
}

int ref_forward(ref_solver *s)
{   /* forward, solver.pyx:413-585 (time interpolation + particle loop + increment) */
    ref_pset *p = s->pset;
    const ref_params *par = s->par;
    int npart = p->npart, i, bnd_j, e;
    if (s->bnd_stats)
        for (i = 0; i < s->mesh->nopen; i++) s->bnd_stats[i] = 0;
    if (s->csec_stats)
        for (i = 0; i < s->ncsec; i++) s->csec_stats[i] = 0;
    e = ref_time_interpolation(s->fset, s->time);
    if (e) { s->err_code = e; s->err_particle = -1; return e; }
    for (i = 0; i < npart; i++) {
        if (p->inactive[i] == 0 && p->tri[i] == -2) {
            e = ref_initialize_particle_state(s, i);
            if (e) { s->err_code = e; s->err_particle = i; return e; }
        }
        if (p->inactive[i] != 0) continue;
        bnd_j = -1;
        if (par->model == 0) e = update_state_test(s, i, &bnd_j);
        else if (par->model == 1) e = update_state_lsm1(s, i, &bnd_j);
        else if (par->model == 2) e = ref_update_state_lsm2(s, i, &bnd_j);
        else e = ref_update_state_lsm3(s, i, &bnd_j);
        if (e) { s->err_code = e; s->err_particle = i; return e; }
        if (par->boundary_conditions) {
            e = compute_bc(s, i, bnd_j);
            if (e) { s->err_code = e; s->err_particle = i; return e; }
        } else {
            if (p->tri[i] == PART_LOC_O) ref_delete_single(p, i);
            if (p->lowerlayer[i] == PART_LOC_B) ref_delete_single(p, i);
            if (p->lowerlayer[i] == PART_LOC_A) ref_delete_single(p, i);
        }
        if (p->compute_depth && p->inactive[i] == 0 && p->dim == 3 && p->depth) {
            double zs;
            e = ref_interp_field_2d(s->fset, s->mesh, s->fset->fid_zs, p->tri[i],
                                    s->fset->nlayers - 1, p->position[i * 3],
                                    p->position[i * 3 + 1], &zs);
            if (e) { s->err_code = e; s->err_particle = i; return e; }
            p->depth[i] = zs - p->position[i * 3 + 2];
        }
        /* control sections, solver.pyx:540-549 (a row deleted above has both
         * positions zeroed by delete_single) */
        for (int c = 0; c < s->ncsec && s->csec_stats; c++) {
            const double *a = p->last_position + i * p->dim, *b = p->position + i * p->dim;
            if (csec_cross(s, c, a[0], a[1], b[0], b[1])) s->csec_stats[c] += 1;
        }
    }
    /* SimuTime.increment, simutime.pyx:77-85 */
    s->time += s->time_step;
    s->iteration += 1;
    return REF_OK;
}

int ref_forward_n(ref_solver *s, int nsteps)
{
    for (int n = 0; n < nsteps; n++) {
        int e = ref_forward(s);
        if (e) return e;
    }
    return REF_OK;
}
