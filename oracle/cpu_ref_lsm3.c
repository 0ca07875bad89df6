/*
 * oracle/cpu_ref_lsm3.c -- LSM-3 of the CPU restatement.  TEST INFRASTRUCTURE
 * ONLY; see cpu_ref.h.  Restates cylag/lsm/update_state_lsm3.pyx (frozen-step
 * coefficients of the (Us, Ur = Up - A2 Us, Xp) system, Cholesky factors,
 * Euler-Maruyama and direct integrators) and init_lsm3_fluctuations.pyx.
 *
 * The coefficient routines are pinned against matrix exponentials (scipy
 * expm of the drift and Van Loan's block exponential for the covariance),
 * the integrators against the reference's one-step closed forms.
 */
#include "cpu_ref.h"
#include <float.h>
#include <math.h>

#define EPSILON 1.e-16
#define EPSILON_LSM2 1.e-8
#define EPSILON_LSM3 1.e-8
#define LSM3_CLIP_STEP 1.e-4
#define GRAV 9.80665
#define C0KOLM 1.2
#define CK(e) do { int _e = (e); if (_e) return _e; } while (0)

int ref_mean_velocity_p3d(ref_solver *s, int k, int ll, double x, double y, double z, double *ux,
                          double *uy, double *uz);

/* ---- frozen-step coefficients ------------------------------------------- */
static int lsm3_bad_rates(double a, double a2, double tl, double dt)
{   /* input checks, update_state_lsm3.pyx:74-81 (and :324-331) */
    if (!isfinite(a) || !isfinite(a2) || !isfinite(tl) || !isfinite(dt) || a < 0. || tl <= 0. || dt < 0.)
        return 1;
    double b = 1. / tl;
    if (!isfinite(b) || !isfinite(a * dt) || !isfinite(b * dt) || !isfinite(a * (1. - a2) * dt)) return 1;
    return 0;
}

static void lsm3_identity(double *phi, double *q, double *force, double *flow)
{
    for (int k = 0; k < 9; k++) { phi[k] = 0.; q[k] = 0.; }
    phi[0] = phi[4] = phi[8] = 1.;
    for (int k = 0; k < 3; k++) { force[k] = 0.; flow[k] = 0.; }
}

static int lsm3_finite(const double *phi, const double *q, const double *force, const double *flow)
{   /* sanity checks, update_state_lsm3.pyx:206-217 */
    for (int k = 0; k < 9; k++)
        if (!isfinite(phi[k]) || !isfinite(q[k])) return 0;
    if (!isfinite(force[1]) || !isfinite(force[2])) return 0;
    for (int k = 0; k < 3; k++)
        if (!isfinite(flow[k])) return 0;
    return 1;
}

static int lsm3_series_coefficients(double a, double a2, double tl, double dt, double *phi, double *q,
                                    double *force, double *flow)
{   /* _lsm3_direct_coefficients, update_state_lsm3.pyx:32-217: 24-term Taylor
     * series at max(a h, b h) <= 1/2, then doublings of the step */
    if (lsm3_bad_rates(a, a2, tl, dt)) return REF_ERR_LSM3_COEF;
    double b = 1. / tl;
    lsm3_identity(phi, q, force, flow);
    if (dt == 0.) return REF_OK;
    double h = dt;
    int nd = 0;
    while (h > 0. && fmax(a * h, b * h) > 0.5) {
        h *= 0.5;
        nd++;
    }
    double ah = a * h, bh = b * h, kh = a * (1. - a2) * h;
    /* second moments of the normalized states (s, p, x, si): ss sp sx pp px xx si pi xi ii */
    double t[10] = {1., 0., 0., 0., 0., 0., 0., 0., 0., 0.};
    double qb[10] = {1., 0., 0., 0., 0., 0., 0., 0., 0., 0.};
    double s = 1., pv = 0., xv = 0., si = 0.;
    double sum_p = 0., sum_x = 0., sum_si = 0., int_p = 0., int_x = 0., int_si = 0.;
    double bterm = 1., phi1 = 1., phi2 = 0.5;
    for (int n = 1; n < 25; n++) {
        double dv = n + 1., u[10];
        u[0] = -2. * bh * t[0] / dv;
        u[1] = (t[0] - (ah + bh) * t[1]) / dv;
        u[2] = (t[1] - bh * t[2]) / dv;
        u[3] = (2. * t[1] - 2. * ah * t[3]) / dv;
        u[4] = (t[2] + t[3] - ah * t[4]) / dv;
        u[5] = 2. * t[4] / dv;
        u[6] = (t[0] - bh * t[6]) / dv;
        u[7] = (t[6] + t[1] - ah * t[7]) / dv;
        u[8] = (t[7] + t[2]) / dv;
        u[9] = 2. * t[6] / dv;
        for (int k = 0; k < 10; k++) {
            t[k] = u[k];
            qb[k] += t[k];
        }
        /* first column of exp(N) */
        double sn = -bh * s / n, pn = (s - ah * pv) / n, xn = pv / n, sin_ = s / n;
        s = sn; pv = pn; xv = xn; si = sin_;
        sum_p += pv;
        sum_x += xv;
        sum_si += si;
        int_p += pv / (n + 1.);
        int_x += xv / (n + 1.);
        int_si += si / (n + 1.);
        bterm *= -ah / n;
        phi1 += bterm / (n + 1.);
        phi2 += bterm / ((n + 1.) * (n + 2.));
    }
    q[0] = h * qb[0];
    q[1] = q[3] = h * kh * qb[1];
    q[2] = q[6] = h * h * (kh * qb[2] + a2 * qb[6]);
    q[4] = h * kh * kh * qb[3];
    q[5] = q[7] = h * h * kh * (kh * qb[4] + a2 * qb[7]);
    q[8] = h * h * h * (kh * kh * qb[5] + 2. * kh * a2 * qb[8] + a2 * a2 * qb[9]);
    phi[0] = exp(-bh);
    phi[3] = kh * sum_p;
    phi[4] = exp(-ah);
    phi[6] = h * (kh * sum_x + a2 * sum_si);
    phi[7] = h * phi1;
    force[1] = phi[7];
    force[2] = h * h * phi2;
    flow[0] = -expm1(-bh);
    flow[1] = kh * bh * int_p;
    flow[2] = h * bh * (kh * int_x + a2 * int_si);
    for (int n = 0; n < nd; n++) {
        /* Q(2h) = Q(h) + Phi(h) Q(h) Phi(h)^T with Phi lower triangular */
        double w[9], qn[9];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) {
                double v = 0.;
                for (int k = 0; k <= r; k++) v += phi[3 * r + k] * q[3 * k + c];
                w[3 * r + c] = v;
            }
        for (int r = 0; r < 3; r++)
            for (int c = 0; c <= r; c++) {
                double v = q[3 * r + c];
                for (int k = 0; k <= c; k++) v += w[3 * r + k] * phi[3 * c + k];
                qn[3 * r + c] = qn[3 * c + r] = v;
            }
        for (int k = 0; k < 9; k++) q[k] = qn[k];
        flow[2] = 2. * flow[2] + phi[6] * flow[0] + phi[7] * flow[1];
        flow[1] = (1. + phi[4]) * flow[1] + phi[3] * flow[0];
        flow[0] = -expm1(-2. * b * h);
        phi[6] = phi[6] * (1. + phi[0]) + phi[7] * phi[3];
        phi[3] *= phi[0] + phi[4];
        force[2] = 2. * force[2] + phi[7] * phi[7];
        phi[7] *= 1. + phi[4];
        force[1] = phi[7];
        h *= 2.;
        phi[0] = exp(-b * h);
        phi[4] = exp(-a * h);
    }
    return lsm3_finite(phi, q, force, flow) ? REF_OK : REF_ERR_LSM3_COEF;
}

static double lsm3_phi1(double z)
{   /* (exp(z) - 1)/z, update_state_lsm3.pyx:219-223 */
    if (z == 0.) return 1.;
    return expm1(z) / z;
}

static double lsm3_exp_divdiff(double x, double y)
{   /* (exp(-x) - exp(-y))/(y - x), update_state_lsm3.pyx:225-227 */
    return exp(-fmin(x, y)) * lsm3_phi1(-fabs(x - y));
}

static double lsm3_phi2(double z)
{   /* (exp(z) - 1 - z)/z^2, update_state_lsm3.pyx:229-234 */
    if (fabs(z) < 0.05)
        return 0.5 + z * (1. / 6. + z * (1. / 24. + z * (1. / 120. + z * (1. / 720. +
               z * (1. / 5040. + z * (1. / 40320. + z / 362880.))))));
    return (expm1(z) - z) / (z * z);
}

static void lsm3_closed_covariance(double a, double b, double a2, double h, double *q)
{   /* _lsm3_closed_covariance, update_state_lsm3.pyx:236-266 */
    double d = a - b, c = a * (1. - a2) / d, bb = (1. - a2) / d, aa = (a2 + c) / b;
    double sb = h * lsm3_phi1(-b * h), sa = h * lsm3_phi1(-a * h);
    double s2b = h * lsm3_phi1(-2. * b * h), s2a = h * lsm3_phi1(-2. * a * h);
    double sab = h * lsm3_phi1(-(a + b) * h);
    q[0] = s2b;
    q[1] = q[3] = c * (s2b - sab);
    q[2] = q[6] = aa * (sb - s2b) - bb * (sb - sab);
    q[4] = c * c * (s2b - 2. * sab + s2a);
    q[5] = q[7] = c * (aa * (sb - s2b - sa + sab) - bb * (sb - sab - sa + s2a));
    q[8] = (aa * aa * (h - 2. * sb + s2b) - 2. * aa * bb * (h - sb - sa + sab) + bb * bb * (h - 2. * sa + s2a));
}

static void lsm3_clipped_covariance(double a, double b, double a2, double h, double *q)
{   /* _lsm3_clipped_covariance, update_state_lsm3.pyx:268-310 */
    double d = a - b, m = 0.5 * (a + b), x = m * h, delta;
    if (x > 1.) delta = fmin(0.5, LSM3_CLIP_STEP * sqrt(x)) / h;
    else delta = fmin(0.5, LSM3_CLIP_STEP / sqrt(x)) * m;
    if (fabs(d) >= delta) {
        lsm3_closed_covariance(a, b, a2, h, q);
    } else {
        double qp[9], qm[9], w;
        lsm3_closed_covariance(m + 0.5 * delta, m - 0.5 * delta, a2, h, qp);
        lsm3_closed_covariance(m - 0.5 * delta, m + 0.5 * delta, a2, h, qm);
        w = 0.5 * (1. + d / delta);
        for (int k = 0; k < 9; k++) q[k] = qm[k] + w * (qp[k] - qm[k]);
    }
    for (int k = 0; k < 3; k++) q[4 * k] = fmax(q[4 * k], 0.);
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < r; c++) {
            double bound = sqrt(q[4 * r] * q[4 * c]);
            q[3 * r + c] = q[3 * c + r] = fmax(-bound, fmin(bound, q[3 * r + c]));
        }
}

static int lsm3_clipped_coefficients(double a, double a2, double tl, double dt, double *phi, double *q,
                                     double *force, double *flow)
{   /* _lsm3_clipped_coefficients, update_state_lsm3.pyx:312-374 */
    if (lsm3_bad_rates(a, a2, tl, dt)) return REF_ERR_LSM3_COEF;
    double b = 1. / tl;
    lsm3_identity(phi, q, force, flow);
    if (dt == 0.) return REF_OK;
    double ah = a * dt, bh = b * dt, k = a * (1. - a2);
    double alpha = exp(-bh), beta = exp(-ah), divdiff = lsm3_exp_divdiff(ah, bh);
    double ib = dt * lsm3_phi1(-bh), ia = dt * lsm3_phi1(-ah);
    phi[0] = alpha;
    phi[3] = k * dt * divdiff;
    phi[4] = beta;
    phi[6] = ib - (1. - a2) * dt * divdiff;
    phi[7] = ia;
    force[1] = ia;
    force[2] = dt * dt * lsm3_phi2(-ah);
    flow[0] = -expm1(-bh);
    flow[1] = k * dt * (lsm3_phi1(-ah) - divdiff);
    flow[2] = (a2 * (dt - ib) + (1. - a2) * (dt - ia - ib + dt * divdiff));
    lsm3_clipped_covariance(a, b, a2, dt, q);
    return lsm3_finite(phi, q, force, flow) ? REF_OK : REF_ERR_LSM3_COEF;
}

static int lsm3_coefficients(double a, double a2, double tl, double dt, int edge_opt, double *phi, double *q,
                             double *force, double *flow)
{   /* _lsm3_coefficients, update_state_lsm3.pyx:376-384 */
    if (edge_opt == 2) return lsm3_clipped_coefficients(a, a2, tl, dt, phi, q, force, flow);
    return lsm3_series_coefficients(a, a2, tl, dt, phi, q, force, flow);
}

static int lsm3_physical_coupling(double a, double a2, double tl, double dt, double *phi)
{   /* _lsm3_physical_coupling, update_state_lsm3.pyx:386-433 */
    double b = 1. / tl, coupling = a - a2 * b, h = dt;
    int nd = 0;
    phi[3] = phi[6] = 0.;
    if (dt == 0. || coupling == 0.) return REF_OK;
    while (h > 0. && fmax(a * h, b * h) > 0.5) {
        h *= 0.5;
        nd++;
    }
    double ah = a * h, bh = b * h, s = 1., pv = 0., sum_p = 0., sum_x = 0., term = 1., phi1 = 1.;
    for (int n = 1; n < 25; n++) {
        double pn = (s - ah * pv) / n, xn = pv / n;
        s *= -bh / n;
        pv = pn;
        sum_p += pv;
        sum_x += xn;
        term *= -ah / n;
        phi1 += term / (n + 1.);
    }
    phi[3] = (coupling * h) * sum_p;
    phi[6] = (coupling * h) * (h * sum_x);
    double decay = h * phi1;
    for (int n = 0; n < nd; n++) {
        double alpha = exp(-b * h), beta = exp(-a * h);
        phi[6] = phi[6] * (1. + alpha) + decay * phi[3];
        phi[3] *= alpha + beta;
        decay *= 1. + beta;
        h *= 2.;
    }
    return (isfinite(phi[3]) && isfinite(phi[6])) ? REF_OK : REF_ERR_LSM3_COEF;
}

static void lsm3_clipped_physical_coupling(double a, double a2, double tl, double dt, double *phi)
{   /* _lsm3_clipped_physical_coupling, update_state_lsm3.pyx:435-447 */
    double b = 1. / tl, ah = a * dt, bh = b * dt;
    phi[3] = phi[6] = 0.;
    if (dt == 0.) return;
    phi[3] = (a - a2 * b) * dt * lsm3_exp_divdiff(ah, bh);
    phi[6] = dt * (lsm3_phi1(-bh) - a2 * lsm3_phi1(-ah) - (1. - a2) * lsm3_exp_divdiff(ah, bh));
}

int ref_lsm3_direct_coefficients(double a1, double tl, double dt, double a2, int edge_opt, double *transition,
                                 double *covariance, double *response)
{   /* lsm3_direct_coefficients, update_state_lsm3.pyx:449-518: transition and
     * covariance in the physical state order (Us, Up, Xp), unit noise */
    double phi[9], q[9], force[3], flow[3];
    if (edge_opt != 1 && edge_opt != 2) return REF_ERR_BAD_SCHEME;
    CK(lsm3_coefficients(a1, a2, tl, dt, edge_opt, phi, q, force, flow));
    if (edge_opt == 2) lsm3_clipped_physical_coupling(a1, a2, tl, dt, phi);
    else CK(lsm3_physical_coupling(a1, a2, tl, dt, phi));
    q[4] += a2 * (2. * q[1] + a2 * q[0]);
    q[1] += a2 * q[0];
    q[3] = q[1];
    q[5] += a2 * q[2];
    q[7] = q[5];
    for (int k = 0; k < 9; k++) { transition[k] = phi[k]; covariance[k] = q[k]; }
    for (int k = 0; k < 3; k++) response[k] = force[k];
    return REF_OK;
}

static int lsm3_cholesky(const double *q, double *lower)
{   /* _lsm3_cholesky, update_state_lsm3.pyx:520-553 */
    for (int k = 0; k < 9; k++) lower[k] = 0.;
    for (int r = 0; r < 3; r++) {
        double pivot = q[3 * r + r];
        for (int c = 0; c < r; c++) {
            double res = q[3 * r + c];
            for (int k = 0; k < c; k++) res -= lower[3 * r + k] * lower[3 * c + k];
            if (lower[3 * c + c] > 0.) {
                lower[3 * r + c] = res / lower[3 * c + c];
            } else {
                double tol = (128. * DBL_EPSILON * sqrt(fabs(q[3 * r + r])) * sqrt(fabs(q[3 * c + c])));
                if (fabs(res) > tol) return REF_ERR_LSM3_CHOL;
            }
            pivot -= lower[3 * r + c] * lower[3 * r + c];
        }
        if (pivot < 0.) {
            if (-pivot > 128. * DBL_EPSILON * fabs(q[3 * r + r])) return REF_ERR_LSM3_CHOL;
            pivot = 0.;
        }
        lower[3 * r + r] = sqrt(pivot);
    }
    return REF_OK;
}

static void lsm3_cholesky_clipped(const double *q, double *lower)
{   /* _lsm3_cholesky_clipped, update_state_lsm3.pyx:555-585 */
    for (int k = 0; k < 9; k++) lower[k] = 0.;
    for (int r = 0; r < 3; r++) {
        double rem = fmax(q[3 * r + r], 0.);
        for (int c = 0; c < r; c++) {
            if (lower[3 * c + c] > 0.) {
                double v = q[3 * r + c];
                for (int k = 0; k < c; k++) v -= lower[3 * r + k] * lower[3 * c + k];
                v /= lower[3 * c + c];
                if (v * v > rem) v = copysign(sqrt(rem), v);
                lower[3 * r + c] = v;
                rem -= v * v;
            }
        }
        lower[3 * r + r] = sqrt(fmax(rem, 0.));
    }
}

/* ---- fields at the particle --------------------------------------------- */
static int lsm3_mean_flow(ref_solver *s, int i, const double *xp, double *u, double *acc)
{   /* mean velocity (pseudo-3D corrected) and mean fluid acceleration, always
     * evaluated for LSM-3 (update_state_lsm3.pyx:702-727, 876-909) */
    ref_pset *p = s->pset;
    ref_fset *f = s->fset;
    int ll = p->dim == 3 ? p->lowerlayer[i] : 0;
    double z = p->dim == 3 ? xp[2] : 0.;
    CK(ref_mean_velocity_p3d(s, p->tri[i], ll, xp[0], xp[1], z, &u[0], &u[1], &u[2]));
    acc[0] = acc[1] = acc[2] = 0.;
    if (f->dim == 2)
        return ref_interp_acceleration_2d(f, s->mesh, p->tri[i], xp[0], xp[1], s->time_step, u[0], u[1],
                                          &acc[0], &acc[1]);
    return ref_interp_acceleration_3d(f, s->mesh, p->tri[i], ll, xp[0], xp[1], xp[2], s->time_step, u[0], u[1],
                                      u[2], &acc[0], &acc[1], &acc[2]);
}

static int lsm3_keps(ref_solver *s, int i, const double *xp, double *k, double *eps)
{
    ref_fset *f = s->fset;
    ref_pset *p = s->pset;
    if (f->dim == 2) {
        CK(ref_interp_field_2d(f, s->mesh, f->fid_turbeng, p->tri[i], -1, xp[0], xp[1], k));
        return ref_interp_field_2d(f, s->mesh, f->fid_dissip, p->tri[i], -1, xp[0], xp[1], eps);
    }
    CK(ref_interp_field_3d(f, s->mesh, f->fid_turbeng, p->tri[i], p->lowerlayer[i], 0, xp[0], xp[1], xp[2], k));
    return ref_interp_field_3d(f, s->mesh, f->fid_dissip, p->tri[i], p->lowerlayer[i], 0, xp[0], xp[1], xp[2],
                               eps);
}

static double lsm3_relax_rate(const ref_pset *p, int i, double unorm)
{   /* a1 = a1c*(1.5/dp)*unorm*Cd with the slip floored (update_state_lsm3.pyx:729-735) */
    double dp = p->diameter[i];
    unorm = fmax(EPSILON_LSM2, unorm);
    double Cd = ref_drag_coefficient(p, unorm, dp);
    return p->a1c * (1.5 / dp) * unorm * Cd;
}

/* ---- integrators -------------------------------------------------------- */
static int euler_maruyama_lsm3(ref_solver *s, int i)
{   /* euler_maruyama_lsm3, update_state_lsm3.pyx:664-820 */
    ref_pset *p = s->pset;
    const ref_params *par = s->par;
    int d = p->dim, dm = par->diffusion_model;
    double *xp = p->position + i * d, *up = p->velocity + i * d, *us = p->us + i * d;
    double dt = s->time_step, sqrtdt = sqrt(dt), u[3], acc[3];
    double nux = par->horizontal_diffusivity, nuz = par->vertical_diffusivity;
    double tlh = par->diffusion_lsm3_tl_horizontal, tlv = par->diffusion_lsm3_tl_vertical;
    double turbeng = 0., epsilon = 0., tml, unstml, diffx, diffy, diffz, xix, xiy, xiz = 0.;
    CK(lsm3_mean_flow(s, i, xp, u, acc));
    double du0 = us[0] - up[0], du1 = us[1] - up[1], unorm;
    if (d == 2) unorm = sqrt(du0 * du0 + du1 * du1);
    else {
        double du2 = us[2] - up[2];
        unorm = sqrt(du0 * du0 + du1 * du1 + du2 * du2);
    }
    double a1 = lsm3_relax_rate(p, i, unorm);
    CK(ref_generate_random_pair(&s->rng, par->rng_method, &xix, &xiy));
    if (d == 3) CK(ref_generate_random_single(&s->rng, par->rng_method, &xiz));
    if (dm == 2 || dm == 3) {
        CK(lsm3_keps(s, i, xp, &turbeng, &epsilon));
        epsilon = fmax(epsilon, EPSILON);
    }
    if (dm == 2 || dm == 3) {
        tml = (turbeng / epsilon) / (0.5 + (3. / 4.) * C0KOLM);
        unstml = 1. / fmax(EPSILON_LSM3, tml);
        diffx = sqrt(C0KOLM * epsilon) * sqrtdt;
    } else {
        unstml = 1. / tlh;
        diffx = sqrt(2. * nux) * sqrtdt / tlh;
    }
    diffy = diffx;
    double dusx = (acc[0] + (u[0] - us[0]) * unstml) * dt + diffx * xix;
    double dusy = (acc[1] + (u[1] - us[1]) * unstml) * dt + diffy * xiy;
    up[0] += (us[0] - up[0]) * dt * a1;
    up[1] += (us[1] - up[1]) * dt * a1;
    us[0] += dusx;
    us[1] += dusy;
    if (p->added_mass_force) {
        up[0] += p->a2c * dusx;
        up[1] += p->a2c * dusy;
    }
    xp[0] += up[0] * dt;
    xp[1] += up[1] * dt;
    if (d == 3) {
        if (dm == 1 || dm == 2) {
            unstml = 1. / tlv;
            diffz = sqrt(2. * nuz) * sqrtdt / tlv;
        } else {
            tml = (turbeng / epsilon) / (0.5 + (3. / 4.) * C0KOLM);
            unstml = 1. / fmax(EPSILON_LSM3, tml);
            diffz = sqrt(C0KOLM * epsilon) * sqrtdt;
        }
        double dusz = (acc[2] + (u[2] - us[2]) * unstml) * dt + diffz * xiz;
        up[2] += (us[2] - up[2]) * dt * a1 - GRAV * p->a0c * dt;
        us[2] += dusz;
        if (p->added_mass_force) up[2] += p->a2c * dusz;
        xp[2] += up[2] * dt;
    }
    return REF_OK;
}

static int di_lsm3(ref_solver *s, int i)
{   /* di_lsm3, update_state_lsm3.pyx:822-995 */
    ref_pset *p = s->pset;
    const ref_params *par = s->par;
    int d = p->dim, dm = par->diffusion_model, eo = par->lsm3_edge_opt, factored = 0;
    double *xp = p->position + i * d, *up = p->velocity + i * d, *us = p->us + i * d;
    double dt = s->time_step, x0[3] = {0., 0., 0.}, u0[3] = {0., 0., 0.}, s0[3] = {0., 0., 0.};
    double u[3], acc[3], phi[9], q[9], force[3], flow[3], lower[9];
    double a2 = p->added_mass_force ? p->a2c : 0.;
    double tlh = par->diffusion_lsm3_tl_horizontal, tlv = par->diffusion_lsm3_tl_vertical;
    double turbeng = 0., epsilon = 0., sigh, sigv = 0.;
    if (dt == 0.) return REF_OK;
    for (int k = 0; k < d; k++) { x0[k] = xp[k]; u0[k] = up[k]; s0[k] = us[k]; }
    CK(lsm3_mean_flow(s, i, x0, u, acc));
    if (dm == 2 || dm == 3) CK(lsm3_keps(s, i, x0, &turbeng, &epsilon));
    double unorm = (s0[0] - u0[0]) * (s0[0] - u0[0]) + (s0[1] - u0[1]) * (s0[1] - u0[1]);
    if (d == 3) unorm += (s0[2] - u0[2]) * (s0[2] - u0[2]);
    double a1 = lsm3_relax_rate(p, i, sqrt(unorm));
    if (dm == 2 || dm == 3) {
        epsilon = fmax(epsilon, EPSILON);
        tlh = fmax(EPSILON_LSM3, (turbeng / epsilon) / (0.5 + (3. / 4.) * C0KOLM));
        sigh = sqrt(C0KOLM * epsilon);
    } else {
        sigh = sqrt(2. * par->horizontal_diffusivity) / tlh;
    }
    if (d == 3) {
        if (dm == 3) { tlv = tlh; sigv = sigh; }
        else sigv = sqrt(2. * par->vertical_diffusivity) / tlv;
    }
    CK(lsm3_coefficients(a1, a2, tlh, dt, eo, phi, q, force, flow));
    double sig = sigh;
    for (int k = 0; k < d; k++) {
        double tlk = tlh;
        if (k == 2) {
            tlk = tlv;
            sig = sigv;
            if (tlv != tlh) {
                CK(lsm3_coefficients(a1, a2, tlv, dt, eo, phi, q, force, flow));
                factored = 0;
            }
        }
        double uk = u[k], Fp = k == 2 ? -GRAV * p->a0c : 0.;
        double ur0 = u0[k] - a2 * s0[k];
        double us_m = (s0[k] * phi[0] + uk * flow[0] + acc[k] * tlk * flow[0]);
        double ur_m = (ur0 * phi[4] + s0[k] * phi[3] + uk * flow[1] + acc[k] * tlk * flow[1] + Fp * force[1]);
        double xp_m = (x0[k] + ur0 * force[1] + s0[k] * phi[6] + uk * flow[2] + acc[k] * tlk * flow[2] +
                       Fp * force[2]);
        if (sig > 0. && q[0] > 0.) {
            double xix, xiy = 0., xiz = 0.;
            if (!factored) {
                if (eo == 2) lsm3_cholesky_clipped(q, lower);
                else CK(lsm3_cholesky(q, lower));
                factored = 1;
            }
            if (lower[4] > 0.) CK(ref_generate_random_pair(&s->rng, par->rng_method, &xix, &xiy));
            else CK(ref_generate_random_single(&s->rng, par->rng_method, &xix));
            if (lower[8] > 0.) CK(ref_generate_random_single(&s->rng, par->rng_method, &xiz));
            us_m += sig * lower[0] * xix;
            ur_m += sig * (lower[3] * xix + lower[4] * xiy);
            xp_m += sig * (lower[6] * xix + lower[7] * xiy + lower[8] * xiz);
        }
        us[k] = us_m;
        up[k] = ur_m + a2 * us_m;
        xp[k] = xp_m;
    }
    return REF_OK;
}

int ref_update_state_lsm3(ref_solver *s, int i, int *bnd_j)
{   /* update_state_lsm3, update_state_lsm3.pyx:587-662 */
    ref_pset *p = s->pset;
    const ref_params *par = s->par;
    int d = p->dim;
    *bnd_j = -1;
    for (int c = 0; c < d; c++) {
        p->last_position[i * d + c] = p->position[i * d + c];
        p->last_velocity[i * d + c] = p->velocity[i * d + c];
    }
    if (par->time_scheme == 1) CK(euler_maruyama_lsm3(s, i));
    else if (par->time_scheme == 5) CK(di_lsm3(s, i));
    else return REF_ERR_BAD_SCHEME;
    p->last_tri[i] = p->tri[i];
    p->tri[i] = ref_xy_localize(s->mesh, p->last_tri[i], p->last_position[i * d], p->last_position[i * d + 1],
                                p->position[i * d], p->position[i * d + 1], 3, bnd_j);
    if (d == 3)
        p->lowerlayer[i] = ref_z_localize(s->fset, s->mesh, p->tri[i], p->position[i * d],
                                          p->position[i * d + 1], p->position[i * d + 2]);
    return REF_OK;
}

int ref_lsm3_fluctuations(ref_solver *s, int i, double *f1, double *f2, double *f3)
{   /* update_lsm3_fluctuations, init_lsm3_fluctuations.pyx:28-90 */
    ref_pset *p = s->pset;
    const ref_params *par = s->par;
    int d = p->dim, dm = par->diffusion_model;
    double *xp = p->position + i * d, sigh, sigv, xix, xiy, xiz;
    *f1 = *f2 = *f3 = 0.;
    if (dm == 2 || dm == 3) {
        double turbeng, epsilon, tml;
        double xq[3] = {xp[0], xp[1], d == 3 ? xp[2] : 0.};
        CK(lsm3_keps(s, i, xq, &turbeng, &epsilon));
        epsilon = fmax(epsilon, EPSILON);
        tml = fmax(EPSILON_LSM3, (turbeng / epsilon) / (0.5 + (3. / 4.) * C0KOLM));
        sigh = sqrt(0.5 * C0KOLM * epsilon * tml);
    } else {
        sigh = sqrt(par->horizontal_diffusivity / par->diffusion_lsm3_tl_horizontal);
    }
    if (sigh > 0.) {
        CK(ref_generate_random_pair(&s->rng, par->rng_method, &xix, &xiy));
        *f1 = sigh * xix;
        *f2 = sigh * xiy;
    }
    if (d == 3) {
        sigv = dm == 3 ? sigh : sqrt(par->vertical_diffusivity / par->diffusion_lsm3_tl_vertical);
        if (sigv > 0.) {
            CK(ref_generate_random_single(&s->rng, par->rng_method, &xiz));
            *f3 = sigv * xiz;
        }
    }
    return REF_OK;
}
