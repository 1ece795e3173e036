/* Oracle of order 1 (single scattering, grouped pencils) — TEST INFRASTRUCTURE ONLY.
 *
 * FP64 restatement of the reference's grouped order-1 solver:
 *   compress        rings -> (cell, pixel) pencils with moments
 *                   (prompt.py:1317-1402, _prompt_numba.py:1014-1230, 2247-2354)
 *   grouped         per OM, pencils -> (tau, r, theta) groups -> one path grid per
 *                   (tau, r) batch -> quadratic-in-time panels
 *                   (_prompt_numba.py:1750-2065, batch core :257-650)
 *   screen          the charge-only coarse pass with parent/child pencils and the
 *                   two-node inverse-radius rule (_prompt_screen.py:96-477)
 * Groups are formed by sorting (key, enumeration order) and summing in that
 * order, which is the order in which the reference's hash rows receive their
 * additions.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define TWO_PI (2.0 * M_PI)
#define NMOM 23

/* ------------------------------------------------------------ compress */

typedef struct {
    int64_t key;
    int64_t seq;
    double cp, sp;   /* ring azimuth cosine/sine of the node (reference recurrence) */
    int32_t element, sub;
} Node;

static int cmp_node(const void* a, const void* b) {
    const Node* x = (const Node*)a;
    const Node* y = (const Node*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return (x->seq > y->seq) - (x->seq < y->seq);
}

/* Equi-angular cube-sphere pixel (_prompt_numba.py:1015-1043). */
static int64_t pixel_of(double sx, double sy, double sz, int n_face) {
    double ax = fabs(sx), ay = fabs(sy), az = fabs(sz), u, v;
    int face;
    if (ax >= ay && ax >= az) { face = sx > 0 ? 0 : 1; u = sy / ax; v = sz / ax; }
    else if (ay >= az) { face = sy > 0 ? 2 : 3; u = sx / ay; v = sz / ay; }
    else { face = sz > 0 ? 4 : 5; u = sx / az; v = sy / az; }
    double quarter = 0.25 * M_PI;
    int i = (int)((atan(u) + quarter) / (2.0 * quarter) * n_face);
    int j = (int)((atan(v) + quarter) / (2.0 * quarter) * n_face);
    if (i >= n_face) i = n_face - 1;
    if (j >= n_face) j = n_face - 1;
    if (i < 0) i = 0;
    if (j < 0) j = 0;
    return ((int64_t)face * n_face + i) * n_face + j;
}

static int64_t cell_key(double px, double py, double pz, double cell, const double* origin,
                        const int64_t* dims) {
    int64_t c[3];
    double p[3] = {px, py, pz};
    for (int k = 0; k < 3; ++k) {
        int64_t v = (int64_t)floor((p[k] - origin[k]) / cell);
        if (v < 0) v = 0;
        if (v > dims[k] - 1) v = dims[k] - 1;
        c[k] = v;
    }
    return (c[0] * dims[1] + c[1]) * dims[2] + c[2];
}

typedef struct {
    const double *start, *dir, *length, *c0, *c2, *cone, *t0, *t1;
} Chords;

static void sub_point(const Chords* e, int i, int ia, int n_a, double* p, double* t) {
    double frac = (ia + 0.5) / n_a;
    for (int k = 0; k < 3; ++k) p[k] = e->start[3 * i + k] + frac * e->length[i] * e->dir[3 * i + k];
    *t = e->t0[i] + frac * (e->t1[i] - e->t0[i]);
}

static int n_sub(double length, double a_step) {
    int n = (int)ceil(length / a_step);
    return n > 1 ? n : 1;
}

static void ring_basis(const double* u, double* e1, double* e2) {
    double h[3];
    if (fabs(u[2]) < 0.9) { h[0] = 0; h[1] = 0; h[2] = 1; } else { h[0] = 1; h[1] = 0; h[2] = 0; }
    e1[0] = u[1] * h[2] - u[2] * h[1];
    e1[1] = u[2] * h[0] - u[0] * h[2];
    e1[2] = u[0] * h[1] - u[1] * h[0];
    double norm = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    for (int k = 0; k < 3; ++k) e1[k] /= norm;
    e2[0] = u[1] * e1[2] - u[2] * e1[1];
    e2[1] = u[2] * e1[0] - u[0] * e1[2];
    e2[2] = u[0] * e1[1] - u[1] * e1[0];
}

/* One pencil's 23 moments from one ring node (compress_subsegments, :1192-1223). */
static void node_moments(const Chords* e, const Node* nd, double a_step, int ring_nodes,
                         double* row) {
    int i = nd->element, ia = nd->sub;
    const double* u = e->dir + 3 * i;
    double e1[3], e2[3];
    ring_basis(u, e1, e2);
    double cc = e->cone[i];
    double cs = sqrt(fmax(1.0 - cc * cc, 0.0));
    int n_a = n_sub(e->length[i], a_step);
    double share = 1.0 / ((double)n_a * ring_nodes);
    double arc2 = pow(TWO_PI * cs / ring_nodes, 2) / 12.0;
    double sub2 = pow(e->length[i] / n_a, 2) / 12.0;
    double w0 = e->c0[i] * share, w2 = e->c2[i] * share;
    double p[3], t;
    sub_point(e, i, ia, n_a, p, &t);
    double cp = nd->cp, sp = nd->sp;
    double sx = cc * u[0] + cs * (cp * e1[0] + sp * e2[0]);
    double sy = cc * u[1] + cs * (cp * e1[1] + sp * e2[1]);
    double sz = cc * u[2] + cs * (cp * e1[2] + sp * e2[2]);
    double px = p[0], py = p[1], pz = p[2];
    row[0] += w0;
    row[1] += w2;
    row[2] += w0 * px;
    row[3] += w0 * py;
    row[4] += w0 * pz;
    row[5] += w0 * t;
    row[6] += w0 * sx;
    row[7] += w0 * sy;
    row[8] += w0 * sz;
    row[9] += w0 * (px * px + py * py + pz * pz + sub2);
    row[10] += w0 * t * t;
    double tx = -sp * e1[0] + cp * e2[0];
    double ty = -sp * e1[1] + cp * e2[1];
    double tz = -sp * e1[2] + cp * e2[2];
    row[11] += w0 * (sx * sx + arc2 * tx * tx);
    row[12] += w0 * (sy * sy + arc2 * ty * ty);
    row[13] += w0 * (sz * sz + arc2 * tz * tz);
    row[14] += w0 * (sx * sy + arc2 * tx * ty);
    row[15] += w0 * (sx * sz + arc2 * tx * tz);
    row[16] += w0 * (sy * sz + arc2 * ty * tz);
    row[17] += w0 * (px * px + sub2 * u[0] * u[0]);
    row[18] += w0 * (py * py + sub2 * u[1] * u[1]);
    row[19] += w0 * (pz * pz + sub2 * u[2] * u[2]);
    row[20] += w0 * (px * py + sub2 * u[0] * u[1]);
    row[21] += w0 * (px * pz + sub2 * u[0] * u[2]);
    row[22] += w0 * (py * pz + sub2 * u[1] * u[2]);
}

typedef struct {
    int64_t count;
    int64_t* keys;
    double* rows;   /* (count, NMOM) */
} RowTable;

static RowTable* g_rows = NULL;

/* Rings of every chord -> moment rows per (cell, pixel) key, sorted by key.
 * Returns the number of keys; oracle_o1_rows copies them out. */
int64_t oracle_o1_compress(int n_el, const double* start, const double* dir, const double* length,
                           const double* c0, const double* c2, const double* cone,
                           const double* t0, const double* t1, double cell_m, int n_face,
                           int ring_nodes, const double* origin, const int64_t* dims) {
    Chords e = {start, dir, length, c0, c2, cone, t0, t1};
    double a_step = 0.5 * cell_m;
    int64_t nodes = 0;
    for (int i = 0; i < n_el; ++i) nodes += (int64_t)n_sub(length[i], a_step) * ring_nodes;
    Node* list = (Node*)malloc(sizeof(Node) * (nodes > 0 ? nodes : 1));
    const double golden = 0.6180339887498949;
    int64_t n_pix = 6LL * n_face * n_face;
    double step_c = cos(TWO_PI / ring_nodes), step_s = sin(TWO_PI / ring_nodes);
    int64_t k = 0;
    for (int i = 0; i < n_el; ++i) {
        const double* u = dir + 3 * i;
        double e1[3], e2[3];
        ring_basis(u, e1, e2);
        double cc = cone[i], cs = sqrt(fmax(1.0 - cc * cc, 0.0));
        int n_a = n_sub(length[i], a_step);
        double offset = fmod(i * golden, 1.0) * TWO_PI / ring_nodes;
        for (int ia = 0; ia < n_a; ++ia) {
            double p[3], t;
            sub_point(&e, i, ia, n_a, p, &t);
            int64_t ck = cell_key(p[0], p[1], p[2], cell_m, origin, dims);
            double cp = cos(offset), sp = sin(offset);
            for (int q = 0; q < ring_nodes; ++q) {
                if (q > 0) {
                    double c2n = cp * step_c - sp * step_s;
                    double s2n = sp * step_c + cp * step_s;
                    cp = c2n;
                    sp = s2n;
                }
                double sx = cc * u[0] + cs * (cp * e1[0] + sp * e2[0]);
                double sy = cc * u[1] + cs * (cp * e1[1] + sp * e2[1]);
                double sz = cc * u[2] + cs * (cp * e1[2] + sp * e2[2]);
                Node* nd = list + k;
                nd->key = ck * n_pix + pixel_of(sx, sy, sz, n_face);
                nd->seq = k;
                nd->cp = cp;
                nd->sp = sp;
                nd->element = i;
                nd->sub = ia;
                ++k;
            }
        }
    }
    qsort(list, nodes, sizeof(Node), cmp_node);
    int64_t unique = 0;
    for (int64_t j = 0; j < nodes; ++j)
        if (j == 0 || list[j].key != list[j - 1].key) ++unique;
    if (g_rows) { free(g_rows->keys); free(g_rows->rows); free(g_rows); }
    g_rows = (RowTable*)calloc(1, sizeof(RowTable));
    g_rows->count = unique;
    g_rows->keys = (int64_t*)malloc(sizeof(int64_t) * (unique > 0 ? unique : 1));
    g_rows->rows = (double*)calloc((size_t)(unique > 0 ? unique : 1) * NMOM, sizeof(double));
    int64_t r = -1;
    for (int64_t j = 0; j < nodes; ++j) {
        if (j == 0 || list[j].key != list[j - 1].key) {
            ++r;
            g_rows->keys[r] = list[j].key;
        }
        node_moments(&e, list + j, a_step, ring_nodes, g_rows->rows + r * NMOM);
    }
    free(list);
    return unique;
}

void oracle_o1_rows(int64_t* keys, double* rows) {
    if (!g_rows) return;
    memcpy(keys, g_rows->keys, sizeof(int64_t) * g_rows->count);
    memcpy(rows, g_rows->rows, sizeof(double) * NMOM * g_rows->count);
}

/* Moment rows -> pencils (finish_pencils, _prompt_numba.py:2247-2354); rows with
 * w0 <= 0 are dropped. Returns the pencil count. */
int64_t oracle_o1_finish(int64_t count, const double* rows, double* position, double* direction,
                         double* w0, double* w2, double* time, double* major, double* var_major,
                         double* var_minor, double* pos_cov) {
    int64_t q = 0;
    for (int64_t p = 0; p < count; ++p) {
        const double* a = rows + p * NMOM;
        double w = a[0];
        if (w <= 0.0) continue;
        double mx = a[2] / w, my = a[3] / w, mz = a[4] / w;
        position[3 * q] = mx; position[3 * q + 1] = my; position[3 * q + 2] = mz;
        time[q] = a[5] / w;
        w0[q] = w;
        w2[q] = a[1];
        double rx = a[6] / w, ry = a[7] / w, rz = a[8] / w;
        double norm = sqrt(rx * rx + ry * ry + rz * rz);
        double dx = rx / norm, dy = ry / norm, dz = rz / norm;
        direction[3 * q] = dx; direction[3 * q + 1] = dy; direction[3 * q + 2] = dz;
        double cxx = a[11] / w - rx * rx, cyy = a[12] / w - ry * ry, czz = a[13] / w - rz * rz;
        double cxy = a[14] / w - rx * ry, cxz = a[15] / w - rx * rz, cyz = a[16] / w - ry * rz;
        double d[3] = {dx, dy, dz}, e1[3], e2[3];
        ring_basis(d, e1, e2);
        double c11 = (e1[0] * (cxx * e1[0] + cxy * e1[1] + cxz * e1[2])
                      + e1[1] * (cxy * e1[0] + cyy * e1[1] + cyz * e1[2])
                      + e1[2] * (cxz * e1[0] + cyz * e1[1] + czz * e1[2]));
        double c22 = (e2[0] * (cxx * e2[0] + cxy * e2[1] + cxz * e2[2])
                      + e2[1] * (cxy * e2[0] + cyy * e2[1] + cyz * e2[2])
                      + e2[2] * (cxz * e2[0] + cyz * e2[1] + czz * e2[2]));
        double c12 = (e1[0] * (cxx * e2[0] + cxy * e2[1] + cxz * e2[2])
                      + e1[1] * (cxy * e2[0] + cyy * e2[1] + cyz * e2[2])
                      + e1[2] * (cxz * e2[0] + cyz * e2[1] + czz * e2[2]));
        double half = 0.5 * (c11 + c22);
        double disc = sqrt(fmax(0.25 * (c11 - c22) * (c11 - c22) + c12 * c12, 0.0));
        var_major[q] = fmax(half + disc, 0.0);
        var_minor[q] = fmax(half - disc, 0.0);
        double angle = 0.5 * atan2(2.0 * c12, c11 - c22);
        double ca = cos(angle), sa = sin(angle);
        for (int k = 0; k < 3; ++k) major[3 * q + k] = ca * e1[k] + sa * e2[k];
        pos_cov[6 * q] = fmax(a[17] / w - mx * mx, 0.0);
        pos_cov[6 * q + 1] = fmax(a[18] / w - my * my, 0.0);
        pos_cov[6 * q + 2] = fmax(a[19] / w - mz * mz, 0.0);
        pos_cov[6 * q + 3] = a[20] / w - mx * my;
        pos_cov[6 * q + 4] = a[21] / w - mx * mz;
        pos_cov[6 * q + 5] = a[22] / w - my * mz;
        ++q;
    }
    return q;
}

/* ----------------------------------------------------------- batch core */

static const double* g_taylor = NULL;   /* (3, 24) 1 / (n! (n + j + 1)) */

static void ifuncs(double z, double ez, double* i0, double* i1, double* i2) {
    if (fabs(z) < 1.0) {
        double a = g_taylor[23], b = g_taylor[24 + 23], c = g_taylor[48 + 23];
        for (int n = 22; n >= 0; --n) {
            a = a * z + g_taylor[n];
            b = b * z + g_taylor[24 + n];
            c = c * z + g_taylor[48 + n];
        }
        *i0 = a; *i1 = b; *i2 = c;
        return;
    }
    *i0 = (ez - 1.0) / z;
    *i1 = (ez * (z - 1.0) + 1.0) / (z * z);
    *i2 = (ez * (z * z - 2.0 * z + 2.0) - 2.0) / (z * z * z);
}

static double quad_integral(double e_start, double alpha, double beta, double gamma, double width,
                            double kappa, double e_end) {
    double z = kappa * width;
    double ez = e_start > 0.0 ? e_end / e_start : exp(z);
    double i0, i1, i2;
    ifuncs(z, ez, &i0, &i1, &i2);
    return e_start * width * (alpha * i0 + width * (beta * i1 + width * gamma * i2));
}

static int find_bin(const double* edges, int nb, double t) {
    if (t < edges[0]) return -1;
    if (t >= edges[nb]) return nb;
    int low = 0, high = nb;
    while (high - low > 1) {
        int middle = (low + high) / 2;
        if (t < edges[middle]) high = middle; else low = middle;
    }
    return low;
}

typedef struct {
    int nl, nb, direct_bins;
    const double *edges, *eedge;   /* eedge (nl, nb+1) */
    double *pre, *acc;             /* pre (nl, 3), acc (nl, nb, 4) */
} Bins;

/* _deposit_panel (_prompt_numba.py:186-253). */
static void deposit_panel(Bins* B, int lam, double t_a, double t_b, double e_a, double e_b,
                          double alpha, double beta, double gamma, double kappa) {
    int nb = B->nb;
    const double* edges = B->edges;
    const double* eedge = B->eedge + (size_t)lam * (nb + 1);
    double* acc = B->acc + (size_t)lam * nb * 4;
    double* pre = B->pre + (size_t)lam * 3;
    double lo = t_a > edges[0] ? t_a : edges[0];
    double hi = t_b < edges[nb] ? t_b : edges[nb];
    if (hi <= lo) return;
    int k_lo = find_bin(edges, nb, lo);
    int k_hi = find_bin(edges, nb, hi);
    if (k_hi >= nb || (k_hi > k_lo && hi <= edges[k_hi])) k_hi -= 1;
    if (k_hi - k_lo + 1 <= B->direct_bins) {
        for (int k = k_lo; k <= k_hi; ++k) {
            double s_lo = lo > edges[k] ? lo : edges[k];
            double s_hi = hi < edges[k + 1] ? hi : edges[k + 1];
            if (s_hi <= s_lo) continue;
            double e_lo = s_lo == t_a ? e_a : eedge[k];
            if (s_lo != t_a && s_lo != edges[k]) e_lo = exp(kappa * s_lo);
            double e_hi = s_hi == t_b ? e_b : eedge[k + 1];
            if (s_hi != t_b && s_hi != edges[k + 1]) e_hi = exp(kappa * s_hi);
            double shift = s_lo - t_a;
            double a_loc = alpha + shift * (beta + shift * gamma);
            double b_loc = beta + 2.0 * gamma * shift;
            acc[4 * k + 3] += quad_integral(e_lo, a_loc, b_loc, gamma, s_hi - s_lo, kappa, e_hi);
        }
        return;
    }
    double big_a = alpha - t_a * (beta - gamma * t_a);
    double big_b = beta - 2.0 * gamma * t_a;
    double big_c = gamma;
    if (t_a < edges[0]) {
        pre[0] += big_a; pre[1] += big_b; pre[2] += big_c;
    } else {
        int k = k_lo;
        acc[4 * k] += big_a; acc[4 * k + 1] += big_b; acc[4 * k + 2] += big_c;
        acc[4 * k + 3] += quad_integral(e_a, alpha, beta, gamma, edges[k + 1] - t_a, kappa,
                                        eedge[k + 1]);
    }
    if (t_b < edges[nb]) {
        int k = find_bin(edges, nb, t_b);
        acc[4 * k] -= big_a; acc[4 * k + 1] -= big_b; acc[4 * k + 2] -= big_c;
        double shift = t_b - t_a;
        double a_loc = alpha + shift * (beta + shift * gamma);
        double b_loc = beta + 2.0 * gamma * shift;
        acc[4 * k + 3] -= quad_integral(e_b, a_loc, b_loc, gamma, edges[k + 1] - t_b, kappa,
                                        eedge[k + 1]);
    }
}

/* _batch_grid (:258-286). */
static int batch_grid(double r, double sigma, double t_emit_rel, double h, double h_tail,
                      double tail_exp, double mut_min, double speed_max, double edge_end,
                      double* s_buf, int capacity) {
    double ext_end = tail_exp / mut_min;
    double ext_win = speed_max * (edge_end - t_emit_rel) - r;
    if (ext_win < 0.0) ext_win = 0.0;
    if (ext_win > ext_end) ext_win = ext_end;
    double u_end = asinh(ext_end / sigma);
    double u_win = asinh(ext_win / sigma);
    int n_in = u_win > 0.0 ? (int)ceil(u_win / h) : 0;
    int n_out = u_end > u_win ? (int)ceil((u_end - u_win) / h_tail) : 0;
    if (n_in + n_out + 1 > capacity) {
        double scale = (capacity - 1.0) / (n_in + n_out);
        n_in = (int)(n_in * scale);
        n_out = capacity - 1 - n_in;
    }
    s_buf[0] = r;
    int n = 1;
    for (int i = 1; i <= n_in; ++i) s_buf[n++] = r + sigma * sinh(u_win * i / n_in);
    for (int i = 1; i <= n_out; ++i)
        s_buf[n++] = r + sigma * sinh(u_win + (u_end - u_win) * i / n_out);
    return n;
}

static double batch_sigma(double r, double sin_min, double tan_half_min, double y_g) {
    return fmax(0.5 * r * sin_min * fmax(y_g, 0.5 * tan_half_min), 1e-9 * r);
}

typedef struct {
    int n_mono, max_power;
    const double* coef;
    const int64_t *ipow, *jpow;
} Mono;

static void cubic_coefficients(const Mono* M, const double* mom, double* a, double* b) {
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0, b0 = 0, b1 = 0, b2 = 0;
    for (int m = 0; m < M->n_mono; ++m) {
        double value = M->coef[m] * mom[m];
        int64_t i = M->ipow[m], j = M->jpow[m];
        if (j == 0) {
            if (i == 0) a0 += value; else if (i == 1) a1 += value;
            else if (i == 2) a2 += value; else a3 += value;
        } else if (j == 1) {
            if (i == 0) b0 += value; else if (i == 1) b1 += value; else b2 += value;
        } else if (j == 2) {
            if (i == 0) { a0 += value; a2 -= value; } else { a1 += value; a3 -= value; }
        } else {
            b0 += value; b2 -= value;
        }
    }
    a[0] = a0; a[1] = a1; a[2] = a2; a[3] = a3; b[0] = b0; b[1] = b1; b[2] = b2;
}

/* _batch_add_moments (:394-447). */
static void batch_add_moments(int n, const double* s_buf, double r, double sin_t, double cos_t,
                              double jac_sin, const double* mom0, const double* mom2,
                              const Mono* M, double g, double* g0, double* g2, int backward) {
    double rs = r * sin_t;
    double y_theta = backward ? 0.0 : sin_t / (1.0 + cos_t);
    double jac = backward ? 0.0 : 2.0 / (r * jac_sin);
    double g1 = 1.0 + g * g, gg = 2.0 * g;
    double norm = (1.0 - g * g) / (4.0 * M_PI);
    int cubic = M->max_power <= 3;
    double A0[4], B0[3], A2[4], B2[3];
    double cpow[32], spow[32];
    if (cubic) {
        cubic_coefficients(M, mom0, A0, B0);
        cubic_coefficients(M, mom2, A2, B2);
    }
    for (int k = 0; k < n; ++k) {
        double c, s, path_jacobian;
        if (backward) {
            double shift = s_buf[k] - r * cos_t;
            double denominator = rs * rs + shift * shift;
            c = (rs * rs - shift * shift) / denominator;
            s = 2.0 * rs * shift / denominator;
            path_jacobian = 2.0 * r / denominator;
        } else {
            double y = (s_buf[k] - r) / rs + y_theta;
            double inv = 1.0 / (1.0 + y * y);
            c = (1.0 - y * y) * inv;
            s = 2.0 * y * inv;
            path_jacobian = jac * inv;
        }
        double a0, a2;
        if (cubic) {
            a0 = A0[0] + c * (A0[1] + c * (A0[2] + c * A0[3]));
            a0 += s * (B0[0] + c * (B0[1] + c * B0[2]));
            a2 = A2[0] + c * (A2[1] + c * (A2[2] + c * A2[3]));
            a2 += s * (B2[0] + c * (B2[1] + c * B2[2]));
        } else {
            cpow[0] = 1.0; spow[0] = 1.0;
            for (int p = 1; p <= M->max_power; ++p) { cpow[p] = cpow[p - 1] * c; spow[p] = spow[p - 1] * s; }
            a0 = 0.0; a2 = 0.0;
            for (int m = 0; m < M->n_mono; ++m) {
                double basis = M->coef[m] * cpow[M->ipow[m]] * spow[M->jpow[m]];
                a0 += basis * mom0[m];
                a2 += basis * mom2[m];
            }
        }
        double den = g1 - gg * c;
        double value = path_jacobian * norm / (den * sqrt(den));
        g0[k] += a0 * value;
        g2[k] += a2 * value;
    }
}

typedef struct {
    int nl;
    const double *mus, *mut, *speed, *s0w, *s2w;
} Optics1;

/* _batch_deposit (:451-551). */
static void batch_deposit(int n, const double* s_buf, const double* g0, const double* g2,
                          double r, double t_emit_rel, double tail_exp, const Optics1* O,
                          Bins* B, double* charge, double* t_buf, double* h_buf,
                          double front_extra, int compute_bins) {
    const double* edges = B->edges;
    int nb = B->nb;
    for (int lam = 0; lam < O->nl; ++lam) {
        if (O->mus[lam] == 0.0) continue;
        double v = O->speed[lam];
        double kappa = -O->mut[lam] * v;
        int last = n - 1;
        double limit = r + tail_exp / O->mut[lam];
        while (last > 2 && s_buf[last - 1] > limit) --last;
        if (last < 2) continue;
        double spread = fmin(front_extra + fabs(0.0 - 0.0 / v) * 0.0, (s_buf[last] - r) / v);
        double front_time = t_emit_rel + r / v;
        double end_time = t_emit_rel + s_buf[last] / v;
        int outside = (front_time - 0.5 * spread >= edges[nb] || end_time + 0.5 * spread <= edges[0]);
        double t_base = outside ? 0.0 : t_emit_rel;
        double e_emit = exp(kappa * t_base);
        double h_scale = O->mus[lam] * v / e_emit;
        double w0 = O->s0w[lam] * h_scale, w2 = O->s2w[lam] * h_scale;
        for (int j = 0; j <= last; ++j) {
            t_buf[j] = t_base + s_buf[j] / v;
            h_buf[j] = w0 * g0[j] + w2 * g2[j];
        }
        double h_front = h_buf[0];
        for (int j = 0; j <= last; ++j) h_buf[j] -= h_front;
        double total = 0.0;
        double e_first = e_emit * exp(-O->mut[lam] * s_buf[0]);
        double e_a = e_first, e_last = e_first;
        int j = 0;
        while (j < last) {
            double t_a, t_b, alpha, beta, gamma, e_b;
            int step;
            if (j + 2 <= last) {
                t_a = t_buf[j];
                double t_m = t_buf[j + 1];
                t_b = t_buf[j + 2];
                double u_m = t_m - t_a, u_b = t_b - t_a;
                double d1 = (h_buf[j + 1] - h_buf[j]) / u_m;
                double d2 = ((h_buf[j + 2] - h_buf[j + 1]) / (t_b - t_m) - d1) / u_b;
                alpha = h_buf[j];
                beta = d1 - d2 * u_m;
                gamma = d2;
                e_b = e_emit * exp(-O->mut[lam] * s_buf[j + 2]);
                step = 2;
            } else {
                double t_p = t_buf[j - 1];
                t_a = t_buf[j];
                t_b = t_buf[j + 1];
                double u_m = t_a - t_p, u_b = t_b - t_p;
                double d1 = (h_buf[j] - h_buf[j - 1]) / u_m;
                double d2 = ((h_buf[j + 1] - h_buf[j]) / (t_b - t_a) - d1) / u_b;
                alpha = h_buf[j];
                beta = d1 - d2 * u_m + 2.0 * d2 * u_m;
                gamma = d2;
                e_b = e_emit * exp(-O->mut[lam] * s_buf[j + 1]);
                step = 1;
            }
            double width = t_b - t_a;
            total += quad_integral(e_a, alpha, beta, gamma, width, kappa, e_b);
            if (compute_bins && !outside)
                deposit_panel(B, lam, t_a, t_b, e_a, e_b, alpha, beta, gamma, kappa);
            e_a = e_b;
            e_last = e_b;
            j += step;
        }
        double t_f = t_buf[0], t_end = t_buf[last];
        double exact = quad_integral(e_first, h_front, 0.0, 0.0, t_end - t_f, kappa, e_last);
        total += exact;
        if (compute_bins && spread > 0.0 && !outside) {
            double t_1 = t_f - 0.5 * spread, t_2 = t_f + 0.5 * spread;
            double e_1 = exp(kappa * t_1), e_2 = exp(kappa * t_2);
            double ramp = quad_integral(e_1, 0.0, 1.0 / spread, 0.0, spread, kappa, e_2);
            double flat = t_end > t_2 ? quad_integral(e_2, 1.0, 0.0, 0.0, t_end - t_2, kappa, e_last)
                                      : 0.0;
            double height = ramp + flat != 0.0 ? exact / (ramp + flat) : 0.0;
            deposit_panel(B, lam, t_1, t_2, e_1, e_2, 0.0, height / spread, 0.0, kappa);
            if (t_end > t_2) deposit_panel(B, lam, t_2, t_end, e_2, e_last, height, 0.0, 0.0, kappa);
        } else if (compute_bins && t_end > t_f && !outside) {
            deposit_panel(B, lam, t_f, t_end, e_first, e_last, h_front, 0.0, 0.0, kappa);
        }
        charge[lam] += total;
    }
}

/* _finalize (:610-632). */
static double finalize(Bins* B, const double* itab, const double* charge, double area,
                       double* bins_out) {
    int nb = B->nb;
    const double* edges = B->edges;
    double total_charge = 0.0;
    for (int lam = 0; lam < B->nl; ++lam) {
        const double* eedge = B->eedge + (size_t)lam * (nb + 1);
        const double* acc = B->acc + (size_t)lam * nb * 4;
        const double* it = itab + (size_t)lam * nb * 3;
        double a = B->pre[3 * lam], b = B->pre[3 * lam + 1], c = B->pre[3 * lam + 2];
        for (int k = 0; k < nb; ++k) {
            double e_k = edges[k];
            double width = edges[k + 1] - e_k;
            double a_loc = a + e_k * (b + c * e_k);
            double b_loc = b + 2.0 * c * e_k;
            double value = eedge[k] * width
                           * (a_loc * it[3 * k] + width * (b_loc * it[3 * k + 1] + width * c * it[3 * k + 2]));
            bins_out[k] += area * (value + acc[4 * k + 3]);
            a += acc[4 * k];
            b += acc[4 * k + 1];
            c += acc[4 * k + 2];
        }
        total_charge += charge[lam];
    }
    return area * total_charge;
}

/* I_k(kappa w) of every (wavelength, bin) and e^{kappa t} at every edge. */
void oracle_o1_tables(const double* taylor, const double* edges, int nb, const double* kappa,
                      int nl, double* eedge, double* itab) {
    g_taylor = taylor;
    for (int lam = 0; lam < nl; ++lam) {
        for (int k = 0; k <= nb; ++k) eedge[lam * (nb + 1) + k] = exp(kappa[lam] * edges[k]);
        for (int k = 0; k < nb; ++k) {
            double z = kappa[lam] * (edges[k + 1] - edges[k]);
            ifuncs(z, exp(z), itab + (lam * nb + k) * 3, itab + (lam * nb + k) * 3 + 1,
                   itab + (lam * nb + k) * 3 + 2);
        }
    }
}

/* ------------------------------------------------------- pencil geometry */

typedef struct {
    int64_t n;
    const double *pos, *dir, *w0, *w2, *t, *major, *vmaj, *vmin, *cov;
    double* prepared;   /* (n, 7) */
} Pencils;

static void prepare_shape(Pencils* P) {
    for (int64_t p = 0; p < P->n; ++p) {
        const double* a = P->major + 3 * p;
        const double* s = P->dir + 3 * p;
        const double* cov = P->cov + 6 * p;
        double bx = s[1] * a[2] - s[2] * a[1];
        double by = s[2] * a[0] - s[0] * a[2];
        double bz = s[0] * a[1] - s[1] * a[0];
        double* q = P->prepared + 7 * p;
        q[0] = bx; q[1] = by; q[2] = bz;
        q[3] = (a[0] * (cov[0] * a[0] + cov[3] * a[1] + cov[4] * a[2])
                + a[1] * (cov[3] * a[0] + cov[1] * a[1] + cov[5] * a[2])
                + a[2] * (cov[4] * a[0] + cov[5] * a[1] + cov[2] * a[2]));
        q[4] = (bx * (cov[0] * bx + cov[3] * by + cov[4] * bz)
                + by * (cov[3] * bx + cov[1] * by + cov[5] * bz)
                + bz * (cov[4] * bx + cov[5] * by + cov[2] * bz));
        q[5] = (a[0] * (cov[0] * bx + cov[3] * by + cov[4] * bz)
                + a[1] * (cov[3] * bx + cov[1] * by + cov[5] * bz)
                + a[2] * (cov[4] * bx + cov[5] * by + cov[2] * bz));
        q[6] = cov[0] + cov[1] + cov[2];
    }
}

typedef struct { int n; const double *tangent, *weighted_sec2; } Gauss;

static double gauss_inverse_mean(double u, double v, double var_u, double var_v, const Gauss* G) {
    double scale = 2.0 * (var_u + var_v) + u * u + v * v;
    if (scale <= 0.0) return 1e300;
    double s0 = 1.0 / sqrt(scale);
    double total = 0.0;
    for (int k = 0; k < G->n; ++k) {
        double s = s0 * G->tangent[k];
        double s2 = s * s;
        double a = 1.0 + 2.0 * s2 * var_u;
        double b = 1.0 + 2.0 * s2 * var_v;
        total += G->weighted_sec2[k] * exp(-s2 * (u * u / a + v * v / b)) / sqrt(a * b);
    }
    return 2.0 / sqrt(M_PI) * s0 * total;
}

/* _pencil_shape_factor_from_ray (:1389-1436): factor, theta_eff, averaged, ok. */
static int shape_factor(const Pencils* P, int64_t p, double r, double theta, double sine,
                        double hx, double hy, double hz, double smoothing, const Gauss* G,
                        double* factor_out, double* theta_out, int* averaged_out) {
    const double* a = P->major + 3 * p;
    const double* q = P->prepared + 7 * p;
    double inv_r2 = 1.0 / (r * r);
    double m11 = P->vmaj[p] + q[3] * inv_r2;
    double m22 = P->vmin[p] + q[4] * inv_r2;
    double m12 = q[5] * inv_r2;
    double upper = fmax(m11, m22) + fabs(m12);
    double biggest = fmax(fmax(fabs(m11), fabs(m22)), fabs(m12));
    int far = biggest < 1e150
              && theta * theta > smoothing * smoothing * fmax(upper, 0.0) * (1.0 + 1e-12);
    double mu1 = 0, mu2 = 0, extent = 0;
    if (!far) {
        double half_tr = 0.5 * (m11 + m22);
        double disc = sqrt(fmax(0.25 * (m11 - m22) * (m11 - m22) + m12 * m12, 0.0));
        mu1 = half_tr + disc;
        mu2 = fmax(half_tr - disc, 1e-16 * mu1 + 1e-300);
        extent = sqrt(fmax(mu1, 0.0));
    }
    double u = hx * a[0] + hy * a[1] + hz * a[2];
    double v = hx * q[0] + hy * q[1] + hz * q[2];
    if (!far && extent > 0.0 && theta < smoothing * extent) {
        double angle = 0.5 * atan2(2.0 * m12, m11 - m22);
        double ca = cos(angle), sa = sin(angle);
        double mean = gauss_inverse_mean(ca * u + sa * v, -sa * u + ca * v, mu1, mu2, G);
        double factor = mean * (sine > 1e-300 ? theta / sine : 1.0);
        *theta_out = factor > 1.0 ? asin(fmin(1.0, 1.0 / factor)) : theta;
        *factor_out = factor;
        *averaged_out = theta < extent;
        return 1;
    }
    if (sine <= 1e-300) { *factor_out = 0.0; *theta_out = theta; *averaged_out = 0; return 0; }
    double d2 = u * u + v * v;
    double factor = 1.0 / sine;
    if (d2 > 0.0) {
        double quad_form = m11 * u * u + 2.0 * m12 * u * v + m22 * v * v;
        factor *= 1.0 + 0.5 * (3.0 * quad_form / d2 - (m11 + m22)) / d2;
    }
    *factor_out = factor;
    *theta_out = theta;
    *averaged_out = 0;
    return 1;
}

/* inverse_radius_rule (_prompt_screen.py:96-137). */
static void inverse_radius_rule(double u, double v, double var_u, double var_v, const Gauss* G,
                                double* out) {
    double scale = 2.0 * (var_u + var_v) + u * u + v * v;
    out[0] = out[1] = out[2] = out[3] = 0.0;
    if (scale <= 0.0) return;
    double s0 = 1.0 / sqrt(scale);
    double inverse = 0.0, radial = 0.0;
    for (int k = 0; k < G->n; ++k) {
        double t = s0 * G->tangent[k];
        double a = 1.0 + 2.0 * t * t * var_u;
        double b = 1.0 + 2.0 * t * t * var_v;
        double mass = G->weighted_sec2[k] * exp(-t * t * (u * u / a + v * v / b)) / sqrt(a * b);
        inverse += mass;
        radial += mass * (var_u / a + var_v / b + u * u / (a * a) + v * v / (b * b));
    }
    double multiplier = 2.0 / sqrt(M_PI) * s0;
    inverse *= multiplier;
    radial *= multiplier;
    if (inverse <= 0.0) return;
    double mean = 1.0 / inverse;
    double second = radial / inverse;
    double variance = second - mean * mean;
    if (variance <= 0.0) { out[0] = inverse; out[1] = mean; out[2] = 0.0; out[3] = mean; return; }
    double third = (u * u + v * v + var_u + var_v) / inverse;
    double other = (third - 2.0 * mean * second + mean * mean * mean) / variance;
    double disc = sqrt((mean - other) * (mean - other) + 4.0 * variance);
    double low = 0.5 * (mean + other - disc);
    double high = 0.5 * (mean + other + disc);
    double low_weight = inverse * (high - mean) / (high - low);
    out[0] = low_weight; out[1] = low; out[2] = inverse - low_weight; out[3] = high;
}

/* pencil_radial_rule (_prompt_screen.py:140-167): two (weight factor, angle) nodes. */
static int radial_rule(const Pencils* P, int64_t p, double r, double theta, double sine, double hx,
                       double hy, double hz, double smoothing, const Gauss* G, double* f1,
                       double* t1, double* f2, double* t2, int* averaged) {
    const double* q = P->prepared + 7 * p;
    const double* a = P->major + 3 * p;
    double m11 = P->vmaj[p] + q[3] / (r * r);
    double m22 = P->vmin[p] + q[4] / (r * r);
    double m12 = q[5] / (r * r);
    double half = 0.5 * (m11 + m22);
    double disc = sqrt(fmax(0.25 * (m11 - m22) * (m11 - m22) + m12 * m12, 0.0));
    double mu1 = half + disc;
    double mu2 = fmax(half - disc, 1e-16 * mu1 + 1e-300);
    double extent = sqrt(fmax(mu1, 0.0));
    if (extent > 0.0 && theta < smoothing * extent) {
        double u = hx * a[0] + hy * a[1] + hz * a[2];
        double v = hx * q[0] + hy * q[1] + hz * q[2];
        if (sine > 0.0) { u *= theta / sine; v *= theta / sine; }
        double angle = 0.5 * atan2(2 * m12, m11 - m22);
        double ca = cos(angle), sa = sin(angle);
        double rule[4];
        inverse_radius_rule(ca * u + sa * v, -sa * u + ca * v, mu1, mu2, G, rule);
        double w0 = rule[0], a0 = rule[1], w1 = rule[2], a1 = rule[3];
        if (a0 > 0.0 && a1 < 0.5 * M_PI && w0 > 0.0 && w1 >= 0.0) {
            *f1 = w0 * a0 / sin(a0); *t1 = a0;
            *f2 = w1 * a1 / sin(a1); *t2 = a1;
            *averaged = theta < extent;
            return 1;
        }
    }
    double factor, angle;
    int ok = shape_factor(P, p, r, theta, sine, hx, hy, hz, smoothing, G, &factor, &angle, averaged);
    *f1 = factor; *t1 = angle; *f2 = 0.0; *t2 = angle;
    return ok;
}

/* ---------------------------------------------------------- grouped core */

typedef struct {
    int64_t key, seq;
    double w0, w2, c1, c3, theta, r, t_emit, tau;
    int averaged;
} Entry;

static int cmp_entry(const void* a, const void* b) {
    const Entry* x = (const Entry*)a;
    const Entry* y = (const Entry*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return (x->seq > y->seq) - (x->seq < y->seq);
}

static int theta_index(double theta, double theta_lo, double log_ratio) {
    if (theta < theta_lo) return 0;
    return 1 + (int)(log(theta / theta_lo) / log_ratio);
}

/* Moments w c1^i c3^j of both fields (_fill_moments_averaged_pair, :1260-1277). */
static void fill_moments(const Mono* M, double weight0, double weight2, double c1, double c3,
                         int averaged, double* out0, double* out2) {
    double cp[32], sp[32];
    cp[0] = 1.0; sp[0] = 1.0;
    double perp2 = fmax(1.0 - c1 * c1, 0.0);
    for (int p = 1; p <= M->max_power; ++p) {
        cp[p] = cp[p - 1] * c1;
        if (averaged) sp[p] = (p % 2 == 1) ? 0.0 : (p >= 2 ? sp[p - 2] : 1.0) * perp2 * (p - 1.0) / p;
        else sp[p] = sp[p - 1] * c3;
    }
    for (int m = 0; m < M->n_mono; ++m) {
        out0[m] = weight0 * cp[M->ipow[m]] * sp[M->jpow[m]];
        out2[m] = weight2 * cp[M->ipow[m]] * sp[M->jpow[m]];
    }
}

typedef struct {
    const double *receivers, *looks, *areas, *origins;
    const int64_t* work;
    int64_t n_work;
    const int64_t *parent, *next;   /* screen hierarchy or NULL */
    Mono mono;
    Optics1 optics;
    Gauss gauss;
    double g, y_g, h, h_tail, tail_exp, smoothing, theta_lo, theta_ratio, tau_step, log_r_step,
        speed_ref;
    int max_nodes, compute_bins, direct_bins, nb;
    const double *edges, *eedge, *itab;
} GroupedArgs;

/* One OM of the grouped (or screen) pass; returns its charge, adds its bins. */
static double grouped_one(const GroupedArgs* A, Pencils* P, int64_t d, double* bins_out,
                          int64_t* groups_out, int64_t* batches_out) {
    const Mono* M = &A->mono;
    int n_mono = M->n_mono;
    int screen = A->parent != NULL;
    double log_ratio = log(1.0 + A->theta_ratio);
    int64_t theta_bins = 2 + (int64_t)(log(M_PI / A->theta_lo) / log_ratio);
    double rx = A->receivers[3 * d], ry = A->receivers[3 * d + 1], rz = A->receivers[3 * d + 2];
    double nx = A->looks[3 * d], ny = A->looks[3 * d + 1], nz = A->looks[3 * d + 2];
    double origin = A->origins[d];
    Entry* list = (Entry*)malloc(sizeof(Entry) * (2 * P->n + 1));
    int64_t count = 0;
    int64_t cursor = 0;
    while (cursor < P->n) {
        int64_t p = cursor;
        cursor += 1;
        const double* pp = P->pos + 3 * p;
        double dx = rx - pp[0], dy = ry - pp[1], dz = rz - pp[2];
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        double hx = dx / r, hy = dy / r, hz = dz / r;
        double t_emit = screen ? 0.0 : P->t[p] - origin;
        double tau = screen ? 0.0 : t_emit + r / A->speed_ref;
        double log_r = log(r);
        int64_t base;
        if (screen) {
            base = ((int64_t)floor(log(r) / A->log_r_step) + (1LL << 20)) * 2 * theta_bins;
        } else {
            int64_t k0 = (int64_t)floor(log_r / A->log_r_step) + (1LL << 20);
            int64_t k1 = (int64_t)floor(tau / A->tau_step) + (1LL << 20);
            if (k1 < 0) k1 = 0;
            if (k1 > (1LL << 21) - 1) k1 = (1LL << 21) - 1;
            base = (k1 * (1LL << 21) + k0) * 2 * theta_bins;
        }
        const double* s = P->dir + 3 * p;
        double c1 = s[0] * nx + s[1] * ny + s[2] * nz;
        double cosine = hx * s[0] + hy * s[1] + hz * s[2];
        double px = hx - cosine * s[0], py = hy - cosine * s[1], pz = hz - cosine * s[2];
        double sine = sqrt(px * px + py * py + pz * pz);
        double c3 = sine > 1e-300 ? (px * nx + py * ny + pz * nz) / sine : 0.0;
        double theta = atan2(sine, cosine);
        if (screen && A->parent[p] == -2) {
            double extent2 = P->vmaj[p] + P->prepared[7 * p + 6] / (r * r);
            if (theta < 0.5 * M_PI && theta * theta < A->smoothing * A->smoothing * extent2) continue;
            cursor = A->next[p];
        }
        int backward = theta > 0.5 * M_PI;
        double bound = 0.0;
        if (!backward) bound = P->vmaj[p] + P->vmin[p] + P->prepared[7 * p + 6] / (r * r);
        double weight, weight2 = 0.0, theta2 = theta;
        int averaged;
        if (backward) {
            weight = 1.0 / r;
            averaged = 0;
        } else if (sine > 1e-300 && bound < 1e-4 * sine * sine) {
            weight = (1.0 / sine) / r;
            averaged = 0;
        } else if (screen) {
            double f1, t1, f2, t2;
            if (!radial_rule(P, p, r, theta, sine, hx, hy, hz, A->smoothing, &A->gauss, &f1, &t1,
                             &f2, &t2, &averaged)) continue;
            weight = f1 / r;
            weight2 = f2 / r;
            theta = t1;
            theta2 = t2;
        } else {
            double factor, theta_eff;
            if (!shape_factor(P, p, r, theta, sine, hx, hy, hz, A->smoothing, &A->gauss, &factor,
                              &theta_eff, &averaged)) continue;
            weight = factor / r;
            theta = theta_eff;
        }
        int nodes = weight2 > 0.0 ? 2 : 1;
        for (int node = 0; node < nodes; ++node) {
            double w = node == 0 ? weight : weight2;
            double th = node == 0 ? theta : theta2;
            Entry* e = list + count++;
            e->key = base + backward * theta_bins + theta_index(th, A->theta_lo, log_ratio);
            e->seq = 2 * p + node;
            e->w0 = w * P->w0[p];
            e->w2 = w * P->w2[p];
            e->c1 = c1;
            e->c3 = c3;
            e->averaged = averaged;
            e->theta = th;
            e->r = r;
            e->t_emit = t_emit;
            e->tau = tau;
        }
    }
    qsort(list, count, sizeof(Entry), cmp_entry);
    int nv = 2 * n_mono + 6;
    int64_t groups = 0;
    for (int64_t j = 0; j < count; ++j)
        if (j == 0 || list[j].key != list[j - 1].key) ++groups;
    double* values = (double*)calloc((size_t)(groups > 0 ? groups : 1) * nv, sizeof(double));
    int64_t* gkeys = (int64_t*)malloc(sizeof(int64_t) * (groups > 0 ? groups : 1));
    double mom0[64], mom2[64];
    int64_t gi = -1;
    for (int64_t j = 0; j < count; ++j) {
        if (j == 0 || list[j].key != list[j - 1].key) { ++gi; gkeys[gi] = list[j].key; }
        double* row = values + gi * nv;
        const Entry* e = list + j;
        fill_moments(M, e->w0, e->w2, e->c1, e->c3, e->averaged, mom0, mom2);
        for (int m = 0; m < n_mono; ++m) { row[m] += mom0[m]; row[n_mono + m] += mom2[m]; }
        double omega = fabs(mom0[0]);
        row[2 * n_mono] += omega;
        row[2 * n_mono + 1] += omega * e->theta;
        row[2 * n_mono + 2] += omega * e->r;
        row[2 * n_mono + 3] += omega * e->t_emit;
        row[2 * n_mono + 4] += omega * e->tau;
        row[2 * n_mono + 5] += omega * e->tau * e->tau;
    }
    free(list);
    int nl = A->optics.nl, nb = A->nb;
    double* pre = (double*)calloc(3 * nl, sizeof(double));
    double* acc = (double*)calloc((size_t)nl * (nb > 0 ? nb : 1) * 4, sizeof(double));
    double* charge = (double*)calloc(nl, sizeof(double));
    Bins B = {nl, nb, A->direct_bins, A->edges, A->eedge, pre, acc};
    double* s_buf = (double*)malloc(sizeof(double) * A->max_nodes);
    double* t_buf = (double*)malloc(sizeof(double) * A->max_nodes);
    double* h_buf = (double*)malloc(sizeof(double) * A->max_nodes);
    double* g0 = (double*)malloc(sizeof(double) * A->max_nodes);
    double* g2 = (double*)malloc(sizeof(double) * A->max_nodes);
    double mut_min = A->optics.mut[0], speed_max = A->optics.speed[0];
    for (int l = 1; l < nl; ++l) {
        mut_min = fmin(mut_min, A->optics.mut[l]);
        speed_max = fmax(speed_max, A->optics.speed[l]);
    }
    int64_t start = 0, batches = 0;
    while (start < groups) {
        int64_t batch_key = gkeys[start] / theta_bins;
        int backward = batch_key % 2 == 1;
        int64_t stop = start;
        double omega_sum = 0, r_sum = 0, t_sum = 0, tau_sum = 0, tau2_sum = 0, theta_small = M_PI;
        while (stop < groups && gkeys[stop] / theta_bins == batch_key) {
            const double* row = values + stop * nv;
            double omega = row[2 * n_mono];
            if (omega > 0.0) {
                omega_sum += omega;
                r_sum += row[2 * n_mono + 2];
                t_sum += row[2 * n_mono + 3];
                tau_sum += row[2 * n_mono + 4];
                tau2_sum += row[2 * n_mono + 5];
                theta_small = fmin(theta_small, row[2 * n_mono + 1] / omega);
            }
            ++stop;
        }
        if (omega_sum > 0.0) {
            double r_b = r_sum / omega_sum;
            double t_b = t_sum / omega_sum;
            double tau_mean = tau_sum / omega_sum;
            double tau_var = fmax(tau2_sum / omega_sum - tau_mean * tau_mean, 0.0);
            theta_small = fmax(theta_small, 1e-9);
            double sigma;
            if (backward)
                sigma = fmax(0.5 * r_b * sin(theta_small) * A->y_g,
                             0.25 * r_b * (1.0 - cos(theta_small)));
            else
                sigma = batch_sigma(r_b, sin(theta_small), tan(0.5 * theta_small), A->y_g);
            int n = screen ? batch_grid(r_b, sigma, 0.0, A->h, A->h, A->tail_exp, mut_min,
                                        speed_max, INFINITY, s_buf, A->max_nodes)
                           : batch_grid(r_b, sigma, t_b, A->h, A->h_tail, A->tail_exp, mut_min,
                                        speed_max, A->edges[nb], s_buf, A->max_nodes);
            for (int k = 0; k < n; ++k) { g0[k] = 0.0; g2[k] = 0.0; }
            for (int64_t q = start; q < stop; ++q) {
                const double* row = values + q * nv;
                double omega = row[2 * n_mono];
                if (omega <= 0.0) continue;
                double theta = fmax(row[2 * n_mono + 1] / omega, 1e-9);
                double sine = sin(theta);
                batch_add_moments(n, s_buf, r_b, sine, cos(theta), sine, row, row + n_mono, M,
                                  A->g, g0, g2, backward);
            }
            batch_deposit(n, s_buf, g0, g2, r_b, t_b, A->tail_exp, &A->optics, &B, charge, t_buf,
                          h_buf, sqrt(12.0 * tau_var), A->compute_bins);
            ++batches;
        }
        start = stop;
    }
    double result;
    if (A->compute_bins) {
        result = finalize(&B, A->itab, charge, A->areas[d], bins_out);
    } else {
        double total = 0.0;
        for (int l = 0; l < nl; ++l) total += charge[l];
        result = A->areas[d] * total;
    }
    *groups_out = groups;
    *batches_out = batches;
    free(values); free(gkeys); free(pre); free(acc); free(charge);
    free(s_buf); free(t_buf); free(h_buf); free(g0); free(g2);
    return result;
}

/* Grouped order 1 of the OMs in `work` (grouped_pencils_order1 / screen_order1).
 * charge (n_recv) and bins (n_recv, nb) receive the listed OMs' values. */
int oracle_o1_grouped(const double* receivers, const double* looks, const double* areas,
                      const double* origins, const int64_t* work, int64_t n_work,
                      int64_t n_pen, const double* pos, const double* dir, const double* w0,
                      const double* w2, const double* t, const double* major,
                      const double* vmaj, const double* vmin, const double* cov,
                      const int64_t* parent, const int64_t* next, const double* tangent,
                      const double* weighted_sec2, int n_gauss, const double* coef,
                      const int64_t* ipow, const int64_t* jpow, int n_mono, int max_power,
                      double g, double y_g, double h, double h_tail, double tail_exp, int nl,
                      const double* mus, const double* mut, const double* speed,
                      const double* s0w, const double* s2w, const double* edges, int nb,
                      const double* eedge, const double* itab, const double* taylor,
                      int direct_bins, int max_nodes, double smoothing, double theta_lo,
                      double theta_ratio, double tau_step, double log_r_step, double speed_ref,
                      int compute_bins, double* charge_out, double* bins_out,
                      int64_t* groups_out, int64_t* batches_out) {
    g_taylor = taylor;
    if (max_power >= 31 || n_mono > 64) return -1;
    Pencils P = {n_pen, pos, dir, w0, w2, t, major, vmaj, vmin, cov, NULL};
    P.prepared = (double*)malloc(sizeof(double) * 7 * (n_pen > 0 ? n_pen : 1));
    prepare_shape(&P);
    GroupedArgs A;
    memset(&A, 0, sizeof(A));
    A.receivers = receivers; A.looks = looks; A.areas = areas; A.origins = origins;
    A.work = work; A.n_work = n_work; A.parent = parent; A.next = next;
    A.mono.n_mono = n_mono; A.mono.max_power = max_power; A.mono.coef = coef;
    A.mono.ipow = ipow; A.mono.jpow = jpow;
    A.optics.nl = nl; A.optics.mus = mus; A.optics.mut = mut; A.optics.speed = speed;
    A.optics.s0w = s0w; A.optics.s2w = s2w;
    A.gauss.n = n_gauss; A.gauss.tangent = tangent; A.gauss.weighted_sec2 = weighted_sec2;
    A.g = g; A.y_g = y_g; A.h = h; A.h_tail = h_tail; A.tail_exp = tail_exp;
    A.smoothing = smoothing; A.theta_lo = theta_lo; A.theta_ratio = theta_ratio;
    A.tau_step = tau_step; A.log_r_step = log_r_step; A.speed_ref = speed_ref;
    A.max_nodes = max_nodes; A.compute_bins = compute_bins; A.direct_bins = direct_bins;
    A.nb = nb; A.edges = edges; A.eedge = eedge; A.itab = itab;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int64_t i = 0; i < n_work; ++i) {
        int64_t d = work[i];
        charge_out[d] = grouped_one(&A, &P, d, compute_bins ? bins_out + d * nb : NULL,
                                    groups_out + d, batches_out + d);
    }
    free(P.prepared);
    return 0;
}
