/*
 * oracle/cpu_ref_lsm23.c -- LSM-2 of the CPU
 * restatement.  TEST INFRASTRUCTURE ONLY; see cpu_ref.h.
 * Restates cylag/lsm/update_state_lsm2.pyx and init_lsm2_fluctuations.pyx.
 */
#include "cpu_ref.h"
#include <math.h>

#define EPSILON 1.e-16
#define EPSILON_LSM2 1.e-8
#define GRAV 9.80665
#define C0KOLM 1.2
#define CK(e) do { int _e = (e); if (_e) return _e; } while (0)

/* shared with cpu_ref.c */
static int mean_vel(ref_solver *s, int k, int ll, double x, double y, double z, double *ux,
                    double *uy, double *uz)
{
    ref_fset *f = s->fset;
    if (f->dim == 2) {
        CK(ref_interp_velocity_2d(f, s->mesh, k, x, y, ux, uy));
        *uz = 0.;
        if (s->pset->dim == 3) return REF_ERR_BAD_SCHEME; /* pseudo-3D LSM2: see lsm2 entry */
    } else {
        CK(ref_interp_velocity_3d(f, s->mesh, k, ll, 0, x, y, z, ux, uy, uz));
    }
    return REF_OK;
}

/* pseudo-3D needs the corrections of cpu_ref.c; exported there under this name */
int ref_mean_velocity_p3d(ref_solver *s, int k, int ll, double x, double y, double z, double *ux,
                          double *uy, double *uz);

static int velocity_at(ref_solver *s, int k, int ll, double x, double y, double z, double *ux,
                       double *uy, double *uz)
{
    if (s->fset->dim == 2 && s->pset->dim == 3)
        return ref_mean_velocity_p3d(s, k, ll, x, y, z, ux, uy, uz);
    return mean_vel(s, k, ll, x, y, z, ux, uy, uz);
}

static int acceleration_at(ref_solver *s, int k, int ll, double x, double y, double z, double ux,
                           double uy, double uz, double *ax, double *ay, double *az)
{   /* "compute dUs/dt for added mass", update_state_lsm2.pyx:180-189 */
    *ax = 0.; *ay = 0.; *az = 0.;
    if (!s->pset->added_mass_force) return REF_OK;
    if (s->fset->dim == 2)
        return ref_interp_acceleration_2d(s->fset, s->mesh, k, x, y, s->time_step, ux, uy, ax, ay);
    return ref_interp_acceleration_3d(s->fset, s->mesh, k, ll, x, y, z, s->time_step, ux, uy, uz,
                                      ax, ay, az);
}

static double relax_rate(ref_solver *s, int i, double unorm)
{   /* unorm clip, Cd, a1 = a1c*(1.5/dp)*Cd*unorm (update_state_lsm2.pyx:192-198) */
    ref_pset *p = s->pset;
    unorm = fmax(EPSILON_LSM2, unorm);
    double dp = p->diameter[i];
    double Cd = ref_drag_coefficient(p, unorm, dp);
    return p->a1c * (1.5 / dp) * Cd * unorm;
}

static int euler_lsm2(ref_solver *s, int i)
{   /* euler_lsm2, update_state_lsm2.pyx:141-217 */
    ref_pset *p = s->pset;
    int d = p->dim;
    double *xp = p->position + i * d, *up = p->velocity + i * d, *us = p->us + i * d;
    double dt = s->time_step, ux, uy, uz, ax, ay, az, unorm;
    int ll = d == 3 ? p->lowerlayer[i] : 0;
    double z = d == 3 ? xp[2] : 0.;
    CK(velocity_at(s, p->tri[i], ll, xp[0], xp[1], z, &ux, &uy, &uz));
    CK(acceleration_at(s, p->tri[i], ll, xp[0], xp[1], z, ux, uy, uz, &ax, &ay, &az));
    if (d == 2) unorm = sqrt(pow(ux - up[0], 2.0) + pow(uy - up[1], 2.0));
    else unorm = sqrt(pow(ux - up[0], 2.0) + pow(uy - up[1], 2.0) + pow(uz - up[2], 2.0));
    double a1 = relax_rate(s, i, unorm), a2c = p->a2c, a0c = p->a0c;
    us[0] = ux;
    us[1] = uy;
    up[0] += (ux - up[0]) * dt * a1 + ax * a2c * dt;
    up[1] += (uy - up[1]) * dt * a1 + ay * a2c * dt;
    xp[0] += up[0] * dt;
    xp[1] += up[1] * dt;
    if (d == 3) {
        us[2] = uz;
        up[2] += (uz - up[2]) * dt * a1 + az * a2c * dt - GRAV * a0c * dt;
        xp[2] += up[2] * dt;
    }
    return REF_OK;
}

/* rk2/rk3/rk4_lsm2, update_state_lsm2.pyx:219-1021 */
static int rk_lsm2(ref_solver *s, int i, int order, int *bnd_out)
{
    ref_pset *p = s->pset;
    const ref_mesh *m = s->mesh;
    int d = p->dim;
    double *xp = p->position + i * d, *up = p->velocity + i * d, *us = p->us + i * d;
    double dt = s->time_step, ux, uy, uz, ax, ay, az, unorm, a1;
    double a0c = p->a0c, a2c = p->a2c;
    double kx[4] = {0}, ky[4] = {0}, kz[4] = {0}, fx[4] = {0}, fy[4] = {0}, fz[4] = {0};
    /* position fraction of stage s+1 applied to k_s, velocity fraction applied to f_s */
    static const double cpos2[1] = {0.5}, cvel2[1] = {0.5};
    static const double cpos3[2] = {1. / 3., 2. / 3.}, cvel3[2] = {1. / 3., 2. / 3.};
    static const double cpos4[3] = {0.5, 0.5, 1.}, cvel4[3] = {0.5, 0.5, 1.};
    const double *cpos = order == 2 ? cpos2 : order == 3 ? cpos3 : cpos4;
    const double *cvel = order == 2 ? cvel2 : order == 3 ? cvel3 : cvel4;
    int tri = p->tri[i], ll = d == 3 ? p->lowerlayer[i] : 0, bnd = -1;
    double xt = xp[0], yt = xp[1], zt = d == 3 ? xp[2] : 0.;
    *bnd_out = -1;
    for (int st = 0; st < order; st++) {
        CK(velocity_at(s, tri, ll, xt, yt, zt, &ux, &uy, &uz));
        CK(acceleration_at(s, tri, ll, xt, yt, zt, ux, uy, uz, &ax, &ay, &az));
        if (st == 0) {
            if (d == 2) unorm = sqrt(pow(ux - up[0], 2.0) + pow(uy - up[1], 2.0));
            else unorm = sqrt(pow(ux - up[0], 2.0) + pow(uy - up[1], 2.0) + pow(uz - up[2], 2.0));
            a1 = relax_rate(s, i, unorm);
            kx[0] = up[0] * dt;
            ky[0] = up[1] * dt;
            fx[0] = (ux - up[0]) * dt * a1 + ax * a2c * dt;
            fy[0] = (uy - up[1]) * dt * a1 + ay * a2c * dt;
            if (d == 3) {
                kz[0] = up[2] * dt;
                fz[0] = (uz - up[2]) * dt * a1 + az * a2c * dt - GRAV * a0c * dt;
            }
        } else {
            double c = cvel[st - 1];
            if (d == 2)
                unorm = sqrt(pow(ux - up[0] - c * fx[st - 1], 2.0) + pow(uy - up[1] - c * fy[st - 1], 2.0));
            else
                unorm = sqrt(pow(ux - up[0] - c * fx[st - 1], 2.0) + pow(uy - up[1] - c * fy[st - 1], 2.0) +
                             pow(uz - up[2] - c * fz[st - 1], 2.0));
            a1 = relax_rate(s, i, unorm);
            kx[st] = (up[0] + c * fx[st - 1]) * dt;
            ky[st] = (up[1] + c * fy[st - 1]) * dt;
            fx[st] = (ux - (up[0] + c * fx[st - 1])) * dt * a1 + ax * a2c * dt;
            fy[st] = (uy - (up[1] + c * fy[st - 1])) * dt * a1 + ay * a2c * dt;
            if (d == 3) {
                kz[st] = (up[2] + c * fz[st - 1]) * dt;
                fz[st] = (uz - (up[2] + c * fz[st - 1])) * dt * a1 + az * a2c * dt - GRAV * a0c * dt;
            }
        }
        if (st == order - 1) break;
        /* position of the next stage */
        double cp = cpos[st];
        xt = xp[0] + cp * kx[st];
        yt = xp[1] + cp * ky[st];
        zt = d == 3 ? xp[2] + cp * kz[st] : 0.;
        p->last_tri[i] = p->tri[i];
        tri = ref_xy_localize(m, p->last_tri[i], p->last_position[i * d],
                              p->last_position[i * d + 1], xt, yt, 3, &bnd);
        if (d == 3) ll = ref_z_localize(s->fset, m, tri, xt, yt, zt);
        if (bnd != -1 || (d == 3 && (ll == -2 || ll == -3))) {
            p->tri[i] = tri;
            if (d == 3) p->lowerlayer[i] = ll;
            us[0] = ux;
            us[1] = uy;
            up[0] += (ux - up[0]) * dt * a1 + ax * a2c * dt;
            up[1] += (uy - up[1]) * dt * a1 + ay * a2c * dt;
            if (d == 3) {
                us[2] = uz;
                up[2] += (uz - up[2]) * dt * a1 + az * a2c * dt - GRAV * a0c * dt;
            }
            xp[0] += up[0] * dt;
            xp[1] += up[1] * dt;
            if (d == 3) xp[2] += up[2] * dt;
            *bnd_out = bnd;
            return REF_OK;
        }
    }
    us[0] = ux;
    us[1] = uy;
    if (d == 3) us[2] = uz;
    if (order == 2) {
        up[0] += fx[1];
        up[1] += fy[1];
        xp[0] += kx[1];
        xp[1] += ky[1];
        if (d == 3) { up[2] += fz[1]; xp[2] += kz[1]; }
    } else if (order == 3) {
        up[0] += (0.25 * fx[0] + 0.75 * fx[2]);
        up[1] += (0.25 * fy[0] + 0.75 * fy[2]);
        xp[0] += (0.25 * kx[0] + 0.75 * kx[2]);
        xp[1] += (0.25 * ky[0] + 0.75 * ky[2]);
        if (d == 3) {
            up[2] += (0.25 * fz[0] + 0.75 * fz[2]);
            xp[2] += (0.25 * kz[0] + 0.75 * kz[2]);
        }
    } else {
        up[0] += (1. / 6.) * (fx[0] + 2. * fx[1] + 2. * fx[2] + fx[3]);
        up[1] += (1. / 6.) * (fy[0] + 2. * fy[1] + 2. * fy[2] + fy[3]);
        xp[0] += (1. / 6.) * (kx[0] + 2. * kx[1] + 2. * kx[2] + kx[3]);
        xp[1] += (1. / 6.) * (ky[0] + 2. * ky[1] + 2. * ky[2] + ky[3]);
        if (d == 3) {
            up[2] += (1. / 6.) * (fz[0] + 2. * fz[1] + 2. * fz[2] + fz[3]);
            xp[2] += (1. / 6.) * (kz[0] + 2. * kz[1] + 2. * kz[2] + kz[3]);
        }
    }
    return REF_OK;
}

static int diffusion_lsm2(ref_solver *s, int i, int tri0, int ll0)
{   /* diffusion_lsm2, update_state_lsm2.pyx:1023-1086 */
    ref_pset *p = s->pset;
    const ref_params *par = s->par;
    int d = p->dim, dm = par->diffusion_model;
    double dt = s->time_step, sqrtdt = sqrt(dt);
    double xix = 0., xiy = 0., xiz = 0., diffx = 0., diffy = 0., diffz = 0.;
    double driftx = 0., drifty = 0., driftz = 0., Kp = 0., dkdx = 0., dkdy = 0., dkdz = 0., z0 = 0.;
    double *xp = p->position + i * d;
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
    xp[0] += driftx + diffx * xix;
    xp[1] += drifty + diffy * xiy;
    if (d == 3) xp[2] += driftz + diffz * xiz;
    return REF_OK;
}

static int keps_at(ref_solver *s, int i, double x, double y, double z, double *k, double *eps)
{   /* turbulent energy and dissipation at the particle (update_state_lsm2.pyx:1164-1176) */
    ref_fset *f = s->fset;
    ref_pset *p = s->pset;
    if (f->dim == 2) {
        CK(ref_interp_field_2d(f, s->mesh, f->fid_turbeng, p->tri[i], -1, x, y, k));
        CK(ref_interp_field_2d(f, s->mesh, f->fid_dissip, p->tri[i], -1, x, y, eps));
    } else {
        CK(ref_interp_field_3d(f, s->mesh, f->fid_turbeng, p->tri[i], p->lowerlayer[i], 0, x, y, z, k));
        CK(ref_interp_field_3d(f, s->mesh, f->fid_dissip, p->tri[i], p->lowerlayer[i], 0, x, y, z, eps));
    }
    return REF_OK;
}

static int euler_maruyama_lsm2_u(ref_solver *s, int i)
{   /* euler_maruyama_lsm2_u, update_state_lsm2.pyx:1088-1230 */
    ref_pset *p = s->pset;
    const ref_params *par = s->par;
    int d = p->dim, dm = par->diffusion_model, nuopt = par->diffusion_lsm2_option;
    double *xp = p->position + i * d, *up = p->velocity + i * d, *us = p->us + i * d;
    double dt = s->time_step, sqrtdt = sqrt(dt), ux, uy, uz, ax, ay, az, unorm;
    double xix = 0., xiy = 0., xiz = 0., diffx = 0., diffy = 0., diffz = 0.;
    double nux = par->horizontal_diffusivity, nuz = par->vertical_diffusivity;
    int ll = d == 3 ? p->lowerlayer[i] : 0;
    double z = d == 3 ? xp[2] : 0.;
    CK(velocity_at(s, p->tri[i], ll, xp[0], xp[1], z, &ux, &uy, &uz));
    CK(acceleration_at(s, p->tri[i], ll, xp[0], xp[1], z, ux, uy, uz, &ax, &ay, &az));
    if (d == 2) unorm = sqrt(pow(ux - up[0], 2.0) + pow(uy - up[1], 2.0));
    else unorm = sqrt(pow(ux - up[0], 2.0) + pow(uy - up[1], 2.0) + pow(uz - up[2], 2.0));
    double a1 = relax_rate(s, i, unorm);
    if (dm != 0) {
        CK(ref_generate_random_pair(&s->rng, par->rng_method, &xix, &xiy));
        if (d == 3) CK(ref_generate_random_single(&s->rng, par->rng_method, &xiz));
    }
    if (dm == 2 || dm == 3) {
        double tk, ep;
        CK(keps_at(s, i, xp[0], xp[1], z, &tk, &ep));
        ep = ep > EPSILON ? ep : EPSILON;
        double tml = (tk / ep) / (0.5 + (3. / 4.) * C0KOLM);
        diffx = a1 * tml * sqrt(C0KOLM * ep) * sqrtdt;
        diffy = diffx;
        if (d == 3) {
            if (dm == 2) diffz = nuopt == 2 ? nuz * sqrtdt : a1 * sqrt(2. * nuz) * sqrtdt;
            else diffz = diffx;
        }
    } else {
        if (nuopt == 2) {
            diffx = nux * sqrtdt; diffy = diffx;
            if (d == 3) diffz = nuz * sqrtdt;
        } else if (nuopt == 3) {
            diffx = a1 * sqrt(2. * nux) * sqrtdt; diffy = diffx;
            if (d == 3) diffz = a1 * sqrt(2. * nuz) * sqrtdt;
        }
    }
    double a2c = p->a2c, a0c = p->a0c;
    us[0] = ux;
    us[1] = uy;
    up[0] += (ux - up[0]) * dt * a1 + ax * a2c * dt + diffx * xix;
    up[1] += (uy - up[1]) * dt * a1 + ay * a2c * dt + diffy * xiy;
    xp[0] += up[0] * dt;
    xp[1] += up[1] * dt;
    if (d == 3) {
        us[2] = uz;
        up[2] += (uz - up[2]) * dt * a1 + az * a2c * dt - GRAV * a0c * dt + diffz * xiz;
        xp[2] += up[2] * dt;
    }
    return REF_OK;
}

int ref_lsm2_direct_coefficients(double a1, double dt, double *beta, double *D1, double *D2,
                                 double *force_x, double *q_uu, double *q_xu, double *q_xx)
{   /* lsm2_direct_coefficients, update_state_lsm2.pyx:1240-1327 */
    double r = a1 * dt, phi1, phi2, term, power2;
    int n;
    if (!isfinite(a1) || !isfinite(dt) || a1 < 0. || dt < 0. || !isfinite(r)) return REF_ERR_LSM2_COEF;
    *beta = exp(-r);
    if (r < 0.1) {
        phi1 = 1.; phi2 = 0.5; term = 1.;
        for (n = 1; n < 13; n++) {
            term *= -r / (n + 1.);
            phi1 += term;
            phi2 += term / (n + 2.);
        }
        *D1 = dt * phi1;
        *D2 = dt * r * phi2;
        *force_x = dt * dt * phi2;
    } else {
        *D1 = -expm1(-r) / a1;
        *D2 = dt - *D1;
        *force_x = *D2 / a1;
    }
    if (r == 0.) *q_uu = dt;
    else *q_uu = dt * (-expm1(-2. * r) / (2. * r));
    *q_xu = 0.5 * *D1 * *D1;
    if (r < 1.) {
        term = 1. / 6.; power2 = 4.;
        double q = 1. / 3.;
        for (n = 1; n < 24; n++) {
            term *= -r / (n + 3.);
            power2 *= 2.;
            q += (power2 - 2.) * term;
        }
        q *= dt * dt * dt;
        *q_xx = q;
    } else *q_xx = ((dt - 2. * *D1 + *q_uu) / a1) / a1;
    return REF_OK;
}

static int di_lsm2(ref_solver *s, int i, int dm, double nux, double nuz, int nuopt)
{   /* di_lsm2, update_state_lsm2.pyx:1329-1543 */
    ref_pset *p = s->pset;
    const ref_params *par = s->par;
    int d = p->dim, k;
    double *xp = p->position + i * d, *up = p->velocity + i * d, *us = p->us + i * d;
    double dt = s->time_step, xp0[3] = {0}, up0[3] = {0}, ux, uy, uz, ax, ay, az, unorm;
    double beta, D1, D2, fxc, quu, qxu, qxx, b2[3];
    for (k = 0; k < d; k++) { xp0[k] = xp[k]; up0[k] = up[k]; }
    int ll = d == 3 ? p->lowerlayer[i] : 0;
    double z0 = d == 3 ? xp0[2] : 0.;
    CK(velocity_at(s, p->tri[i], ll, xp0[0], xp0[1], z0, &ux, &uy, &uz));
    CK(acceleration_at(s, p->tri[i], ll, xp0[0], xp0[1], z0, ux, uy, uz, &ax, &ay, &az));
    if (d == 2) unorm = sqrt(pow(ux - up0[0], 2.0) + pow(uy - up0[1], 2.0));
    else unorm = sqrt(pow(ux - up0[0], 2.0) + pow(uy - up0[1], 2.0) + pow(uz - up0[2], 2.0));
    double a1 = relax_rate(s, i, unorm);
    CK(ref_lsm2_direct_coefficients(a1, dt, &beta, &D1, &D2, &fxc, &quu, &qxu, &qxx));
    double um_beta = -expm1(-a1 * dt);
    if (dm == 0) { b2[0] = b2[1] = b2[2] = 0.; }
    else if (dm == 2 || dm == 3) {
        double tk, ep;
        CK(keps_at(s, i, xp0[0], xp0[1], z0, &tk, &ep));
        ep = fmax(ep, EPSILON);
        double tml = (tk / ep) / (0.5 + (3. / 4.) * C0KOLM);
        b2[0] = a1 * a1 * C0KOLM * ep * pow(tml, 2.0);
        b2[1] = b2[0];
        if (dm == 2) b2[2] = nuopt == 2 ? nuz * nuz : 2. * nuz * a1 * a1;
        else b2[2] = b2[0];
    } else {
        if (nuopt == 2) { b2[0] = nux * nux; b2[1] = b2[0]; b2[2] = nuz * nuz; }
        else if (nuopt == 3) { b2[0] = 2. * nux * a1 * a1; b2[1] = b2[0]; b2[2] = 2. * nuz * a1 * a1; }
        else { b2[0] = b2[1] = b2[2] = 0.; }
    }
    for (k = 0; k < d; k++) {
        double uk = k == 0 ? ux : k == 1 ? uy : uz;
        double Fp = k == 0 ? p->a2c * ax : k == 1 ? p->a2c * ay : p->a2c * az - GRAV * p->a0c;
        double up_mean = up0[k] * beta + uk * um_beta + Fp * D1;
        double xp_mean = xp0[k] + up0[k] * D1 + uk * D2 + Fp * fxc;
        if (b2[k] == 0. || dt == 0.) {
            us[k] = uk; up[k] = up_mean; xp[k] = xp_mean;
            continue;
        }
        double var_gp = b2[k] * quu, cov = b2[k] * qxu, var_gx = b2[k] * qxx;
        double Lpp = sqrt(fmax(0., var_gp));
        double Lxp = Lpp > 0. ? cov / Lpp : 0.;
        double Lxx = sqrt(fmax(0., var_gx - pow(Lxp, 2.0)));
        double xix, xiy;
        CK(ref_generate_random_pair(&s->rng, par->rng_method, &xix, &xiy));
        double Gp = Lpp * xix, Gx = Lxp * xix + Lxx * xiy;
        us[k] = uk;
        up[k] = up_mean + Gp;
        xp[k] = xp_mean + Gx;
    }
    return REF_OK;
}

int ref_update_state_lsm2(ref_solver *s, int i, int *bnd_j)
{   /* update_state_lsm2, update_state_lsm2.pyx:30-139 */
    ref_pset *p = s->pset;
    const ref_params *par = s->par;
    int d = p->dim, c, ts = par->time_scheme, dm = par->diffusion_model;
    int nuopt = par->diffusion_lsm2_option, tri0 = p->tri[i], ll0 = 0;
    *bnd_j = -1;
    for (c = 0; c < d; c++) {
        p->last_position[i * d + c] = p->position[i * d + c];
        p->last_velocity[i * d + c] = p->velocity[i * d + c];
    }
    if (d == 3) ll0 = p->lowerlayer[i];
    if (nuopt == 1 || (nuopt >= 2 && dm == 0)) {
        if (ts == 1) CK(euler_lsm2(s, i));
        else if (ts >= 2 && ts <= 4) CK(rk_lsm2(s, i, ts, bnd_j));
        else if (ts == 5) CK(di_lsm2(s, i, 0, 0., 0., 2));
        else return REF_ERR_BAD_SCHEME;
        if (dm != 0) CK(diffusion_lsm2(s, i, tri0, ll0));
    } else if (nuopt >= 2 && dm != 0) {
        if (ts == 1) CK(euler_maruyama_lsm2_u(s, i));
        else if (ts == 5)
            CK(di_lsm2(s, i, dm, par->horizontal_diffusivity, par->vertical_diffusivity, nuopt));
        else return REF_ERR_BAD_SCHEME;
    } else return REF_ERR_BAD_SCHEME;
    p->last_tri[i] = p->tri[i];
    p->tri[i] = ref_xy_localize(s->mesh, p->last_tri[i], p->last_position[i * d],
                                p->last_position[i * d + 1], p->position[i * d],
                                p->position[i * d + 1], 3, bnd_j);
    if (d == 3)
        p->lowerlayer[i] = ref_z_localize(s->fset, s->mesh, p->tri[i], p->position[i * d],
                                          p->position[i * d + 1], p->position[i * d + 2]);
    return REF_OK;
}

int ref_lsm2_fluctuations(ref_solver *s, int i, double *f1, double *f2, double *f3)
{   /* update_lsm2_fluctuations, init_lsm2_fluctuations.pyx:28-129 */
    ref_pset *p = s->pset;
    const ref_params *par = s->par;
    double *xp = p->position + i * p->dim;
    double b0, b1, bz, xix, xiy, xiz;
    *f1 = *f2 = *f3 = 0.;
    if (par->diffusion_model == 0) return REF_OK;
    double unorm = EPSILON_LSM2;
    double Cd = ref_drag_coefficient(p, unorm, p->diameter[i]);
    double a1 = p->a1c * (1.5 / p->diameter[i]) * Cd * unorm;
    if (a1 == 0.) {
        if (par->diffusion_lsm2_option == 2) {
            int dm = par->diffusion_model;
            if (((dm != 2 && dm != 3) && par->horizontal_diffusivity != 0.) ||
                (p->dim == 3 && dm != 3 && par->vertical_diffusivity != 0.))
                return REF_ERR_LSM2_COEF;
        }
        return REF_OK;
    }
    if (par->diffusion_model == 2 || par->diffusion_model == 3) {
        double tk, ep, z = p->dim == 3 ? xp[2] : 0.;
        CK(keps_at(s, i, xp[0], xp[1], z, &tk, &ep));
        ep = fmax(ep, EPSILON);
        double tml = (tk / ep) / (0.5 + (3. / 4.) * C0KOLM);
        b0 = a1 * C0KOLM * ep * pow(tml, 2.0);
        b1 = b0;
        if (par->diffusion_model == 2) {
            if (par->diffusion_lsm2_option == 2) bz = pow(par->vertical_diffusivity, 2.0) / a1;
            else bz = 2. * par->vertical_diffusivity * a1;
        } else bz = b0;
    } else {
        if (par->diffusion_lsm2_option == 2) {
            b0 = pow(par->horizontal_diffusivity, 2.0) / a1;
            b1 = b0;
            bz = pow(par->vertical_diffusivity, 2.0) / a1;
        } else {
            b0 = 2. * par->horizontal_diffusivity * a1;
            b1 = b0;
            bz = 2. * par->vertical_diffusivity * a1;
        }
    }
    if (b0 > 0.) {
        CK(ref_generate_random_pair(&s->rng, par->rng_method, &xix, &xiy));
        *f1 = sqrt(0.5 * b0) * xix;
        *f2 = sqrt(0.5 * b1) * xiy;
    }
    if (p->dim == 3 && bz > 0.) {
        CK(ref_generate_random_single(&s->rng, par->rng_method, &xiz));
        *f3 = sqrt(0.5 * bz) * xiz;
    }
    return REF_OK;
}
