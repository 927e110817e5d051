/*
 * maddy_oracle.c -- CPU restatement of MADDY's per-step Langevin hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the
 * smoke() entry point and bench.py's cpu_baseline leg load it.  The product
 * path (mt_amd/, libmt_engine.so) never links, calls or falls back to it.
 *
 * Compiled twice (oracle/Makefile): REAL=double -> liboracle_f64.so (the FP64
 * oracle every GPU result is judged against) and REAL=float ->
 * liboracle_f32.so (same precision as the reference, used for trajectory
 * parity).  Built without -ffast-math.
 *
 * Each function cites the reference code (paths relative to the reference
 * repository root) whose behaviour it restates.  The bonded terms are written
 * in attachment-point form rather than the reference's expanded trig: every
 * bonded gap is g = (r_i + a_i) - (r_j + b_j), where a_i, b_j are body-fixed
 * offsets rotated by R = Rz(psi) Ry(theta) Rx(phi) (the rotation whose
 * entries src/compute_cuda.cu:337-352 and :89-96 expand), the translational
 * gradient of |g|^2/2 is g and the angular one is u_q . (a_i x g) with
 * u_phi = Rz Ry e_x, u_theta = Rz e_y, u_psi = e_z.
 *
 * Parity pinning of each function: see tests/test_oracle_*.py (closed forms,
 * finite differences, Newton's third law, lattice known answers, published
 * RNG recurrences and the reference's own ran2.h compiled under oracle/_ref).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REAL
#define REAL double
#endif
typedef REAL real;

/* src/mt.h:23-61 */
#define O_RMT 8.12
#define O_RMON 2.0
#define O_LJ_CUTOFF 6.0
#define O_ZERO 999999
#define O_ANGLE_CUTOFF 1.0
#define O_PAIR_CUTOFF 2.5
#define O_KB 0.0019872041
#define O_MAXLON 8
#define O_MAXLAT 16

/* Parameters (subset of src/parameters.h:246-318 that the hot path reads). */
typedef struct {
    float dt, temp, gammaR, gammaTheta, varR, varTheta, alpha, freeze_temp;
    float C, B_psi, B_fi, B_theta, psi_0, fi_0, theta0_gtp, theta0_gdp;
    float A_long, D_long, A_lat, D_lat, seam_coeff;
    int32_t barrier;
    float a_barr_long, r_barr_long, w_barr_long, a_barr_lat, r_barr_lat, w_barr_lat;
    int32_t lj_on;
    float ljscale, ljsigma6, ljpairscutoff;
    int32_t ljpairsupdatefreq;
    int32_t is_wall;
    float rep_h, rep_r, rep_eps, rep_leftborder;
    int32_t is_assembly;
    int32_t tea_on;
    float tea_a;
    int32_t tea_epsilon_freq, tea_capricious;
    float tea_epsmax;
    int32_t lj_regularize;   /* REGULARIZATION, src/compute_cuda.cu:482-486 */
} or_params;

/* Topology in CSR-batched form: trajectory t owns global beads
 * [off[t], off[t+1]); list entries are LOCAL indices with the reference's
 * sign encoding (src/preparator.cpp:346-351, src/compute_cuda.cu:587-592,
 * :646-658).  Per-monomer arrays of the reference (harmonic, fixed,
 * mon_type) are replicated per bead. */
typedef struct {
    int32_t ntraj;
    int32_t max_lj;
    const int64_t *off;
    const int32_t *harm_count, *harm; /* one slot per bead */
    int32_t *lon_count, *lon;         /* O_MAXLON per bead */
    int32_t *lat_count, *lat;         /* O_MAXLAT per bead */
    int32_t *lj_count, *lj;           /* max_lj per bead */
    const uint8_t *fixed, *extra;
    const int32_t *mon_type, *gtp, *on_tub;
    const float *zs;                  /* wall height per trajectory */
    const float *box;                 /* periodic cubic box side per trajectory, 0 = open */
} or_topo;

typedef struct { float x, y, z, fi, theta, psi, w; } or_coord; /* src/mt.h:63-71 */

/* Minimum image in a cubic box of side L (L == 0: open).  The reference has
 * open boundaries only; periodic boxes are this build's extension for the
 * free-dimer batches. */
static void min_image(real d[3], real L)
{
    if (L > 0)
        for (int k = 0; k < 3; k++) d[k] -= L * rint(d[k] / L);
}

/* ---------------------------------------------------------------- frames */
typedef struct {
    real r[3];
    real n[3];      /* R e_z */
    real ap1[3];    /* R p1_def */
    real ap2[3];    /* R p2_def */
    real uphi[3];   /* Rz Ry e_x */
    real uth[3];    /* Rz e_y */
    real fi, theta, psi;
} frame;

static void lateral_offsets(real p1[3], real p2[3])
{
    /* src/mt.h:44-50 */
    const real c = cos(2.0 * M_PI / 13.0), s = sin(2.0 * M_PI / 13.0);
    p1[0] = 0.5 * O_RMT * (c - 1.0); p1[1] = 0.5 * O_RMT * s;  p1[2] = -3.0 * O_RMON / 13.0;
    p2[0] = 0.5 * O_RMT * (c - 1.0); p2[1] = -0.5 * O_RMT * s; p2[2] = 3.0 * O_RMON / 13.0;
}

static void rot_apply(real cf, real sf, real ct, real st, real cp, real sp, const real v[3], real o[3])
{
    /* R = Rz(psi) Ry(theta) Rx(phi); rows as expanded at src/compute_cuda.cu:337-352 */
    o[0] = cp * ct * v[0] + (cp * st * sf - sp * cf) * v[1] + (cp * st * cf + sp * sf) * v[2];
    o[1] = sp * ct * v[0] + (sp * st * sf + cp * cf) * v[1] + (sp * st * cf - cp * sf) * v[2];
    o[2] = -st * v[0] + ct * sf * v[1] + ct * cf * v[2];
}

static void make_frame(const or_coord *c, frame *f)
{
    real p1[3], p2[3], ez[3] = {0, 0, 1};
    lateral_offsets(p1, p2);
    f->r[0] = c->x; f->r[1] = c->y; f->r[2] = c->z;
    f->fi = c->fi; f->theta = c->theta; f->psi = c->psi;
    const real cf = cos((real)c->fi), sf = sin((real)c->fi);
    const real ct = cos((real)c->theta), st = sin((real)c->theta);
    const real cp = cos((real)c->psi), sp = sin((real)c->psi);
    rot_apply(cf, sf, ct, st, cp, sp, ez, f->n);
    rot_apply(cf, sf, ct, st, cp, sp, p1, f->ap1);
    rot_apply(cf, sf, ct, st, cp, sp, p2, f->ap2);
    f->uphi[0] = cp * ct; f->uphi[1] = sp * ct; f->uphi[2] = -st;
    f->uth[0] = -sp; f->uth[1] = cp; f->uth[2] = 0;
}

/* Gap g = (r_i + a_i) - (r_j + b_j) and gradient of |g|^2/2 wrt bead i's
 * six coordinates (x, y, z, fi, theta, psi). */
static real gap_grad(const frame *fi_, const real a[3], const frame *fj, const real b[3], real L, real g[3],
                     real grad[6])
{
    real d[3] = {fi_->r[0] - fj->r[0], fi_->r[1] - fj->r[1], fi_->r[2] - fj->r[2]};
    min_image(d, L);
    for (int k = 0; k < 3; k++) g[k] = d[k] + (a[k] - b[k]);
    const real t0 = a[1] * g[2] - a[2] * g[1];
    const real t1 = a[2] * g[0] - a[0] * g[2];
    const real t2 = a[0] * g[1] - a[1] * g[0];
    grad[0] = g[0]; grad[1] = g[1]; grad[2] = g[2];
    grad[3] = fi_->uphi[0] * t0 + fi_->uphi[1] * t1 + fi_->uphi[2] * t2;
    grad[4] = fi_->uth[0] * t0 + fi_->uth[1] * t1;
    grad[5] = t2;
    return sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
}

/* Harmonic / longitudinal attachment: R_MON > 0 couples i's +n end to j's +n
 * end pointing back (src/compute_cuda.cu:89-96): d = r_j - r_i - R(n_i+n_j),
 * i.e. a_i = R n_i, b_j = -R n_j, g = -d. */
static real gap_axial(const frame *fi_, const frame *fj, real rmon, real L, real g[3], real grad[6])
{
    real a[3], b[3];
    for (int k = 0; k < 3; k++) { a[k] = rmon * fi_->n[k]; b[k] = -rmon * fj->n[k]; }
    return gap_grad(fi_, a, fj, b, L, g, grad);
}

/* Lateral: g = r_i - r_j + R_i p2 - R_j p1 (src/compute_cuda.cu:337-352);
 * swapped==1 exchanges p1 and p2 (entries <= 0, :309-324). */
static real gap_lateral(const frame *fi_, const frame *fj, int swapped, real L, real g[3], real grad[6])
{
    const real *a = swapped ? fi_->ap1 : fi_->ap2;
    const real *b = swapped ? fj->ap2 : fj->ap1;
    return gap_grad(fi_, a, fj, b, L, g, grad);
}

/* src/compute_cuda.cu:16-30 */
static real dmorse(real D, real a, real x) { return 2 * a * D * (1 - exp(-a * x)) * exp(-a * x); }
static real morse_en(real D, real a, real x) { return D * (1 - exp(-a * x)) * (1 - exp(-a * x)) - D; }
static real dbarr(real a, real r, real w, real x) { return -a * exp(-(x - r) * (x - r) / (2 * w * w)) * (x - r) / (w * w); }
static real barr(real a, real r, real w, real x) { return a * exp(-(x - r) * (x - r) / (2 * w * w)); }

static int traj_of(const or_topo *t, int64_t b)
{
    int lo = 0, hi = t->ntraj - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (t->off[mid] <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

static void decode_axial(int32_t e, int *j, real *rmon)
{
    /* src/compute_cuda.cu:72-78, :192-198: negative -> +r_mon */
    if (e < 0) { *rmon = O_RMON; *j = -e; }
    else { *rmon = -O_RMON; *j = e; }
}

static void decode_lateral(int32_t e, int *j, int *swapped)
{
    /* src/compute_cuda.cu:307-324 */
    if (e <= 0) { *swapped = 1; *j = -e; }
    else { *swapped = 0; *j = e; }
    if (*j == O_ZERO) *j = 0;
}

/* ---------------------------------------------------------------- forces */
/* Restates compute_kernel, src/compute_cuda.cu:32-525.  out: 7 doubles per
 * bead in Coord order (x, y, z, fi, theta, psi, w=0).  Extra beads get 0
 * (the reference leaves d_f untouched there, and d_f is zero between steps,
 * :963-972). */
void or_forces(const or_params *p, const or_topo *t, const or_coord *rc, int64_t nb, double *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nb; b++) {
        double *o = out + 7 * b;
        for (int k = 0; k < 7; k++) o[k] = 0;
        if (t->extra[b]) continue;
        const int tr = traj_of(t, b);
        const int64_t base = t->off[tr];
        const real L = t->box ? t->box[tr] : 0;
        frame fi_, fj;
        make_frame(&rc[b], &fi_);
        real F[6] = {0, 0, 0, 0, 0, 0}; /* x y z fi theta psi */
        real g[3], gr[6];
        /* harmonic: :70-182 */
        for (int k = 0; k < t->harm_count[b]; k++) {
            int j; real rm;
            decode_axial(t->harm[b], &j, &rm);
            make_frame(&rc[base + j], &fj);
            const real dr = gap_axial(&fi_, &fj, rm, L, g, gr);
            for (int q = 0; q < 6; q++) F[q] -= (real)p->C * gr[q];
            if (dr < O_ANGLE_CUTOFF) {
                const real psiji = fj.psi - fi_.psi, thetaji = fj.theta - fi_.theta, fiji = fj.fi - fi_.fi;
                const real th0 = (t->gtp[b] == 1) ? p->theta0_gtp : p->theta0_gdp;
                if (rm > 0) {
                    F[5] += p->B_psi * sin(psiji - p->psi_0);
                    F[3] += p->B_fi * sin(fiji - p->fi_0);
                    F[4] += p->B_theta * sin(thetaji - th0);
                } else {
                    F[5] -= p->B_psi * sin(-psiji - p->psi_0);
                    F[3] -= p->B_fi * sin(-fiji - p->fi_0);
                    F[4] -= p->B_theta * sin(-thetaji - th0);
                }
            }
        }
        /* longitudinal: :189-299 */
        for (int k = 0; k < t->lon_count[b]; k++) {
            int j; real rm;
            decode_axial(t->lon[b * O_MAXLON + k], &j, &rm);
            make_frame(&rc[base + j], &fj);
            const real dr = gap_axial(&fi_, &fj, rm, L, g, gr);
            real dUdr = (dr == 0) ? 0 : dmorse(p->D_long, p->A_long, dr) / dr;
            if (p->barrier && t->on_tub[b] == 0 && t->on_tub[base + j] == 0)
                if (dr != 0) dUdr += dbarr(p->a_barr_long, p->r_barr_long, p->w_barr_long, dr) / dr;
            for (int q = 0; q < 6; q++) F[q] -= dUdr * gr[q];
            if (dr < O_ANGLE_CUTOFF) {
                const real psiji = fj.psi - fi_.psi, thetaji = fj.theta - fi_.theta, fiji = fj.fi - fi_.fi;
                const int64_t last = (rc[b].z > rc[base + j].z) ? b : base + j; /* :282 */
                const real th0 = (t->gtp[last] == 1) ? p->theta0_gtp : p->theta0_gdp;
                if (rm > 0) {
                    F[5] += p->B_psi * sin(psiji - p->psi_0);
                    F[3] += p->B_fi * sin(fiji - p->fi_0);
                    F[4] += p->B_theta * sin(thetaji - th0);
                } else {
                    F[5] -= p->B_psi * sin(-psiji - p->psi_0);
                    F[3] -= p->B_fi * sin(-fiji - p->fi_0);
                    F[4] -= p->B_theta * sin(-thetaji - th0);
                }
            }
        }
        /* lateral: :304-466 */
        for (int k = 0; k < t->lat_count[b]; k++) {
            int j, sw;
            decode_lateral(t->lat[b * O_MAXLAT + k], &j, &sw);
            make_frame(&rc[base + j], &fj);
            const real dr = gap_lateral(&fi_, &fj, sw, L, g, gr);
            real dUdr;
            if (dr == 0) dUdr = 0;
            else if (t->mon_type[b] != t->mon_type[base + j]) dUdr = dmorse(p->D_lat / p->seam_coeff, p->A_lat, dr) / dr;
            else dUdr = dmorse(p->D_lat, p->A_lat, dr) / dr;
            if (p->barrier && t->on_tub[b] == 0 && t->on_tub[base + j] == 0)
                if (dr != 0) dUdr += dbarr(p->a_barr_lat, p->r_barr_lat, p->w_barr_lat, dr) / dr;
            for (int q = 0; q < 6; q++) F[q] -= dUdr * gr[q];
        }
        /* LJ: :470-495; with lj_regularize, the REGULARIZATION cap of :482-486
         * (r_mon = 2 nm, src/mt.h:24) */
        if (p->lj_on) {
            for (int k = 0; k < t->lj_count[b]; k++) {
                const int j = t->lj[b * t->max_lj + k];
                const or_coord *c = &rc[base + j];
                real d[3] = {(real)rc[b].x - c->x, (real)rc[b].y - c->y, (real)rc[b].z - c->z};
                min_image(d, L);
                const real dx = d[0], dy = d[1], dz = d[2];
                const real dr = sqrt(dx * dx + dy * dy + dz * dz);
                if (dr < O_LJ_CUTOFF) {
                    real df = 6.0 / pow(dr, 8);
                    const real s6 = (real)p->ljscale * (real)p->ljsigma6;
                    if (p->lj_regularize && s6 * df * dr > (real)p->gammaR * O_RMON / p->dt)
                        df = (real)p->gammaR * O_RMON / (p->dt * dr * s6);
                    const real s = s6 * df;
                    F[0] += s * dx; F[1] += s * dy; F[2] += s * dz;
                }
            }
        }
        /* walls: :497-517 */
        if (p->is_wall) {
            const real z = rc[b].z, lb = p->rep_leftborder, zt = (real)t->zs[tr] + lb;
            if (z < lb) F[2] += p->rep_eps * fabs(z - lb);
            else if (z > zt) F[2] += -p->rep_eps * fabs(z - zt);
            const real rad2 = (real)rc[b].x * rc[b].x + (real)rc[b].y * rc[b].y;
            if (rad2 > (real)p->rep_r * p->rep_r) {
                const real coeff = -p->rep_eps * (sqrt(rad2) - p->rep_r);
                F[0] += rc[b].x / sqrt(rad2) * coeff;
                F[1] += rc[b].y / sqrt(rad2) * coeff;
            }
        }
        o[0] = F[0]; o[1] = F[1]; o[2] = F[2];
        o[3] = F[3]; o[4] = F[4]; o[5] = F[5];
    }
}

/* --------------------------------------------------------------- energies */
/* Restates energy_kernel, src/compute_cuda.cu:676-911.  out: 7 doubles per
 * bead: U_harm, U_long, U_lat, U_psi, U_fi, U_teta, U_lj (Energies order,
 * src/mt.h:92-100), each halved (:902-908). */
void or_energies(const or_params *p, const or_topo *t, const or_coord *rc, int64_t nb, double *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nb; b++) {
        double *o = out + 7 * b;
        for (int k = 0; k < 7; k++) o[k] = 0;
        if (t->extra[b]) continue;
        const int tr = traj_of(t, b);
        const int64_t base = t->off[tr];
        const real L = t->box ? t->box[tr] : 0;
        frame fi_, fj;
        make_frame(&rc[b], &fi_);
        real U_harm = 0, U_long = 0, U_lat = 0, U_psi = 0, U_fi = 0, U_teta = 0, U_lj = 0;
        real g[3], gr[6];
        for (int k = 0; k < t->harm_count[b]; k++) {
            int j; real rm;
            decode_axial(t->harm[b], &j, &rm);
            make_frame(&rc[base + j], &fj);
            const real dr = gap_axial(&fi_, &fj, rm, L, g, gr);
            U_harm += (p->C / 2) * dr * dr;
            if (dr < O_ANGLE_CUTOFF) {
                const real psiij = fi_.psi - fj.psi, fiij = fi_.fi - fj.fi, thetaji = fj.theta - fi_.theta;
                U_psi += p->B_psi * (1 - cos(psiij - p->psi_0));
                U_fi += p->B_fi * (1 - cos(fiij - p->fi_0));
                U_teta += p->B_theta * (1 - cos(thetaji - ((t->gtp[b] == 1) ? p->theta0_gtp : p->theta0_gdp)));
            }
        }
        for (int k = 0; k < t->lon_count[b]; k++) {
            int j; real rm;
            decode_axial(t->lon[b * O_MAXLON + k], &j, &rm);
            make_frame(&rc[base + j], &fj);
            const real dr = gap_axial(&fi_, &fj, rm, L, g, gr);
            U_long += morse_en(p->D_long, p->A_long, dr);
            if (p->barrier && t->on_tub[b] == 0 && t->on_tub[base + j] == 0)
                U_long += barr(p->a_barr_long, p->r_barr_long, p->w_barr_long, dr);
            if (dr < O_ANGLE_CUTOFF) {
                const real psiij = fi_.psi - fj.psi, fiij = fi_.fi - fj.fi, thetaij = fi_.theta - fj.theta;
                const int64_t last = (rc[b].z > rc[base + j].z) ? b : base + j;
                const real th0 = (t->gtp[last] == 1) ? p->theta0_gtp : p->theta0_gdp;
                U_psi += p->B_psi * (1 - cos(psiij - p->psi_0));
                U_fi += p->B_fi * (1 - cos(fiij - p->fi_0));
                U_teta += p->B_theta * (1 - cos(thetaij - th0));
            }
        }
        for (int k = 0; k < t->lat_count[b]; k++) {
            int j, sw;
            decode_lateral(t->lat[b * O_MAXLAT + k], &j, &sw);
            make_frame(&rc[base + j], &fj);
            const real dr = gap_lateral(&fi_, &fj, sw, L, g, gr);
            if (t->mon_type[b] != t->mon_type[base + j]) U_lat += morse_en(p->D_lat / p->seam_coeff, p->A_lat, dr);
            else U_lat += morse_en(p->D_lat, p->A_lat, dr);
            if (p->barrier && t->on_tub[b] == 0 && t->on_tub[base + j] == 0)
                U_lat += barr(p->a_barr_lat, p->r_barr_lat, p->w_barr_lat, dr);
        }
        if (p->lj_on) {
            for (int k = 0; k < t->lj_count[b]; k++) {
                const int j = t->lj[b * t->max_lj + k];
                const or_coord *c = &rc[base + j];
                real d[3] = {(real)rc[b].x - c->x, (real)rc[b].y - c->y, (real)rc[b].z - c->z};
                min_image(d, L);
                const real dr = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                if (dr < O_LJ_CUTOFF) U_lj += (real)p->ljscale * (real)p->ljsigma6 / pow(dr, 6);
            }
        }
        o[0] = U_harm / 2; o[1] = U_long / 2; o[2] = U_lat / 2; o[3] = U_psi / 2;
        o[4] = U_fi / 2; o[5] = U_teta / 2; o[6] = U_lj / 2;
    }
}

/* ------------------------------------------------------------ pair lists */
/* Restates pairs_kernel, src/compute_cuda.cu:527-674 (brute force over all
 * j of the trajectory, in increasing j).  Returns the number of beads whose
 * candidate count exceeded a list's capacity (the reference does not check,
 * :589-592); such entries are dropped here. */
int64_t or_pairs(const or_topo *t, const or_coord *rc, int64_t nb)
{
    int64_t overflow = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : overflow)
    for (int64_t b = 0; b < nb; b++) {
        t->lon_count[b] = 0;
        t->lat_count[b] = 0;
        if (t->extra[b]) continue;
        const int tr = traj_of(t, b);
        const int64_t base = t->off[tr];
        const int n = (int)(t->off[tr + 1] - base);
        const int i = (int)(b - base);
        const int32_t h = t->harm[b];
        const real L = t->box ? t->box[tr] : 0;
        const real rm = (h < 0) ? -O_RMON : O_RMON; /* :548-551 */
        const int hpart = abs(h);
        frame fi_, fj;
        make_frame(&rc[b], &fi_);
        real g[3], gr[6];
        int ovf = 0;
        for (int j = 0; j < n; j++) {
            if (t->mon_type[b] != t->mon_type[base + j] && hpart != j) {
                make_frame(&rc[base + j], &fj);
                const real dr = gap_axial(&fi_, &fj, rm, L, g, gr);
                if (dr < O_PAIR_CUTOFF) {
                    if (t->lon_count[b] < O_MAXLON) t->lon[b * O_MAXLON + t->lon_count[b]++] = (h < 0) ? j : -j;
                    else ovf = 1;
                }
            }
        }
        for (int j = 0; j < n; j++) {
            if (i != j && hpart != j) {
                make_frame(&rc[base + j], &fj);
                for (int ind = 0; ind < 2; ind++) {
                    /* ind 0: p1<->p2 swapped, stored negative (:612-619, :646-651) */
                    const real dr = gap_lateral(&fi_, &fj, ind == 0, L, g, gr);
                    if (dr < O_PAIR_CUTOFF) {
                        int32_t e = (ind == 0) ? (j != 0 ? -j : -O_ZERO) : (j != 0 ? j : O_ZERO);
                        if (t->lat_count[b] < O_MAXLAT) t->lat[b * O_MAXLAT + t->lat_count[b]++] = e;
                        else ovf = 1;
                    }
                }
            }
        }
        overflow += ovf;
    }
    return overflow;
}

/* Restates LJ_kernel, src/compute_cuda.cu:913-940. */
int64_t or_lj_list(const or_params *p, const or_topo *t, const or_coord *rc, int64_t nb)
{
    int64_t overflow = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : overflow)
    for (int64_t b = 0; b < nb; b++) {
        t->lj_count[b] = 0;
        if (t->extra[b]) continue;
        const int tr = traj_of(t, b);
        const int64_t base = t->off[tr];
        const int n = (int)(t->off[tr + 1] - base);
        const int i = (int)(b - base);
        int ovf = 0;
        const real L = t->box ? t->box[tr] : 0;
        for (int j = 0; j < n; j++) {
            real d[3] = {(real)rc[b].x - rc[base + j].x, (real)rc[b].y - rc[base + j].y,
                         (real)rc[b].z - rc[base + j].z};
            min_image(d, L);
            const real dr = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            if (dr < (real)p->ljpairscutoff && i != j) {
                if (t->lj_count[b] < t->max_lj) t->lj[b * t->max_lj + t->lj_count[b]++] = j;
                else ovf = 1;
            }
        }
        overflow += ovf;
    }
    return overflow;
}

/* ------------------------------------------------------------ HybridTaus */
/* src/HybridTaus.cu:53-98 (GPU Gems 3, ch. 37) and src/ht.cuh:8-10. */
static uint32_t taus_step(uint32_t *z, int s1, int s2, int s3, uint32_t m)
{
    uint32_t b = (((*z << s1) ^ *z) >> s2);
    return *z = (((*z & m) << s3) ^ b);
}
static uint32_t lcg_step(uint32_t *z) { return *z = 1664525u * *z + 1013904223u; }
static uint32_t hybrid_taus(uint32_t s[4])
{
    return taus_step(&s[0], 13, 19, 12, 4294967294u) ^ taus_step(&s[1], 2, 25, 4, 4294967288u) ^
           taus_step(&s[2], 3, 11, 17, 4294967280u) ^ lcg_step(&s[3]);
}
static float uint_to_float(uint32_t u)
{
    uint32_t it = 0x3f800000u | (0x007fffffu & u);
    float f;
    memcpy(&f, &it, 4);
    f -= 1.0f;
    return f == 0 ? 1.0e-8f : f;
}
/* One rforce() call: advances the state four times, returns four normals. */
void or_rforce(uint32_t s[4], float out[4])
{
    float r = sqrtf(-2.0f * logf(uint_to_float(hybrid_taus(s))));
    float th = (float)(2.0f * M_PI * uint_to_float(hybrid_taus(s)));
    out[0] = r * sinf(th);
    out[1] = r * cosf(th);
    r = sqrtf(-2.0f * logf(uint_to_float(hybrid_taus(s))));
    th = (float)(2.0f * M_PI * uint_to_float(hybrid_taus(s)));
    out[2] = r * sinf(th);
    out[3] = r * cosf(th);
}
/* Raw integer draws (for bit-exact stream checks). */
void or_taus_draws(uint32_t s[4], int64_t n, uint32_t *out)
{
    for (int64_t k = 0; k < n; k++) out[k] = hybrid_taus(s);
}

/* Numerical Recipes ran2 (src/ran2.h:18-56) with state held in a struct
 * instead of function statics; a positive seed skips the table fill, so the
 * shuffle table starts all-zero with iy = 0 and idum2 = 123456789. */
typedef struct { int32_t idum, idum2, iy, iv[32]; } or_ran2;
static double ran2_next(or_ran2 *st)
{
    const int32_t IM1 = 2147483563, IM2 = 2147483399, IA1 = 40014, IA2 = 40692, IQ1 = 53668, IQ2 = 52774,
                  IR1 = 12211, IR2 = 3791, NTAB = 32, IMM1 = IM1 - 1, NDIV = 1 + IMM1 / NTAB;
    const double AM = 1.0 / IM1, RNMX = 1.0 - 1.2e-7;
    int32_t j, k;
    if (st->idum <= 0) {
        if (-(st->idum) < 1) st->idum = 1;
        else st->idum = -(st->idum);
        st->idum2 = st->idum;
        for (j = NTAB + 7; j >= 0; j--) {
            k = st->idum / IQ1;
            st->idum = IA1 * (st->idum - k * IQ1) - k * IR1;
            if (st->idum < 0) st->idum += IM1;
            if (j < NTAB) st->iv[j] = st->idum;
        }
        st->iy = st->iv[0];
    }
    k = st->idum / IQ1;
    st->idum = IA1 * (st->idum - k * IQ1) - k * IR1;
    if (st->idum < 0) st->idum += IM1;
    k = st->idum2 / IQ2;
    st->idum2 = IA2 * (st->idum2 - k * IQ2) - k * IR2;
    if (st->idum2 < 0) st->idum2 += IM2;
    j = st->iy / NDIV;
    st->iy = st->iv[j] - st->idum2;
    st->iv[j] = st->idum;
    if (st->iy < 1) st->iy += IMM1;
    double t = AM * st->iy;
    return t > RNMX ? RNMX : t;
}
/* generateSeeds, src/HybridTaus.cu:32-48: each word (unsigned)(ran2*UINT_MAX),
 * redrawn while < 128.  out: np uint4 = 4*np words. */
void or_generate_seeds(int32_t seed, int64_t np, uint32_t *out)
{
    or_ran2 st;
    memset(&st, 0, sizeof st);
    st.idum = seed;
    st.idum2 = 123456789;
    for (int64_t i = 0; i < 4 * np; i++) {
        uint32_t v;
        do { v = (uint32_t)(ran2_next(&st) * 4294967295.0); } while (v < 128);
        out[i] = v;
    }
}

/* ------------------------------------------------------------ integrator */
/* Restates integrate_kernel, src/compute_cuda.cu:943-975.  Noise per bead:
 * 6 floats (xyz from rforce(p), angles from rforce(p + B)); when noise is
 * NULL the HybridTaus states `seeds` (2*nb uint4) are advanced instead,
 * only for mobile beads (:949-953). */
void or_integrate(const or_params *p, const or_topo *t, or_coord *rc, const double *F, int64_t nb,
                  const float *noise, uint32_t *seeds)
{
    const real dtg = (real)p->dt / p->gammaR;
    const real dta = (real)p->dt / ((real)p->gammaTheta * p->alpha);
    const real dtt = (real)p->dt / p->gammaTheta;
    const real vta = (real)p->varTheta * sqrt((real)p->freeze_temp / p->alpha);
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nb; b++) {
        if (t->fixed[b] || t->extra[b]) continue;
        float rx[4], ra[4];
        if (noise) {
            for (int k = 0; k < 3; k++) { rx[k] = noise[6 * b + k]; ra[k] = noise[6 * b + 3 + k]; }
        } else {
            or_rforce(seeds + 4 * b, rx);
            or_rforce(seeds + 4 * (b + nb), ra);
        }
        const double *f = F + 7 * b;
        or_coord *c = &rc[b];
        c->x = (float)((real)c->x + dtg * (real)f[0] + (real)p->varR * rx[0]);
        c->y = (float)((real)c->y + dtg * (real)f[1] + (real)p->varR * rx[1]);
        c->z = (float)((real)c->z + dtg * (real)f[2] + (real)p->varR * rx[2]);
        c->fi = (float)((real)c->fi + dta * (real)f[3] + vta * ra[0]);
        c->psi = (float)((real)c->psi + dta * (real)f[5] + vta * ra[1]);
        c->theta = (float)((real)c->theta + dtt * (real)f[4] + (real)p->varTheta * ra[2]);
    }
}

/* -------------------------------------------------------------------- TEA */
/* D_ij/D_ii for the RPY tensor, src/bdhitea_kernel.cu:38-58.  e = unit
 * vector from i to j, r = distance. */
static void rpy_hat(real a, const real e[3], real r, real D[9])
{
    const real ra = r / a;
    real crr, cii;
    if (ra > 2) { crr = 0.75 / ra * (1 - 2 / ra / ra); cii = 0.75 / ra * (1 + 2.0 / 3.0 / ra / ra); }
    else { crr = 3 * ra / 32; cii = 1 - 9 * ra / 32; }
    for (int x = 0; x < 3; x++)
        for (int y = 0; y < 3; y++) D[3 * x + y] = e[x] * e[y] * crr + (x == y ? cii : 0);
}

/* integrateTea_epsilon_unlisted, src/bdhitea_kernel.cu:61-101: per bead
 * ci[3] = row sums of squares of D_hat_ij over j, eps = sum of all 9 entries. */
void or_tea_eps(const or_params *p, const or_topo *t, const or_coord *rc, int64_t nb, double *ci, double *eps)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t b = 0; b < nb; b++) {
        const int tr = traj_of(t, b);
        const int64_t base = t->off[tr], end = t->off[tr + 1];
        real s[4] = {0, 0, 0, 0};
        const real L = t->box ? t->box[tr] : 0;
        for (int64_t j = base; j < end; j++) {
            if (j == b || t->extra[j] || t->extra[b]) continue;
            real e[3] = {(real)rc[j].x - rc[b].x, (real)rc[j].y - rc[b].y, (real)rc[j].z - rc[b].z};
            min_image(e, L);
            const real r = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
            for (int k = 0; k < 3; k++) e[k] /= r;
            real D[9];
            rpy_hat(p->tea_a, e, r, D);
            for (int x = 0; x < 3; x++) s[x] += D[3 * x] * D[3 * x] + D[3 * x + 1] * D[3 * x + 1] + D[3 * x + 2] * D[3 * x + 2];
            s[3] += D[0] + 2 * D[1] + 2 * D[2] + D[4] + 2 * D[5] + D[8];
        }
        ci[3 * b] = s[0]; ci[3 * b + 1] = s[1]; ci[3 * b + 2] = s[2];
        eps[b] = s[3];
    }
}

/* updateTea, src/bdhitea.cu:57-118: per-trajectory epsilon and beta in FP64.
 * Returns 0, or 1 + trajectory index of the first capricious abort
 * (eps > 1, eps > epsmax, or |a| < 1e-7 with tea_a > 0). */
int or_tea_beta(const or_params *p, const or_topo *t, const double *eps_bead, double *eps_t, double *beta_t)
{
    int err = 0;
    for (int tr = 0; tr < t->ntraj; tr++) {
        double s = 0;
        int64_t nn = 0;
        for (int64_t b = t->off[tr]; b < t->off[tr + 1]; b++) {
            s += eps_bead[b];
            if (!t->extra[b]) nn++;
        }
        double eps = s / (3. * nn * (3. * nn - 3.));
        if (eps > 1.0) {
            if (p->tea_capricious && !err) err = 1 + tr;
            eps = 1.0;
        }
        if (eps > p->tea_epsmax && !err) err = 1 + tr;
        const double a = (3. * nn - 1.) * eps * eps - (3. * nn - 2.) * eps;
        double beta;
        if (fabs(a) < 1e-7) {
            beta = 0.5;
            if (p->tea_capricious && p->tea_a > 0.0f && !err) err = 1 + tr;
        } else {
            beta = (1. - sqrt(1. - a)) / a;
        }
        eps_t[tr] = eps;
        beta_t[tr] = (float)beta;
    }
    return err;
}

/* One TEA integration step, src/bdhitea_kernel.cu:16-36 (prepare) and
 * :148-213 (kernel), given forces F (7 per bead), per-bead ci (3), beta per
 * trajectory and noise (6 per bead: xyz noise = rforce(p), angle noise =
 * rforce(p + B)); TEA draws both for every bead (:22, :194).  When noise is
 * NULL the HybridTaus states are advanced. */
void or_tea_integrate(const or_params *p, const or_topo *t, or_coord *rc, const double *F, int64_t nb,
                      const double *ci_bead, const double *beta_t, const float *noise, uint32_t *seeds)
{
    real *R = (real *)malloc(sizeof(real) * 3 * nb);
    float *ang = (float *)malloc(sizeof(float) * 3 * nb);
    real *pos = (real *)malloc(sizeof(real) * 3 * nb);
    const real var = sqrtf(2.0f * (float)O_KB * p->temp * p->gammaR / p->dt);
    for (int64_t b = 0; b < nb; b++) {
        float rx[4], ra[4];
        if (noise) {
            for (int k = 0; k < 3; k++) { rx[k] = noise[6 * b + k]; ra[k] = noise[6 * b + 3 + k]; }
        } else {
            or_rforce(seeds + 4 * b, rx);
            or_rforce(seeds + 4 * (b + nb), ra);
        }
        for (int k = 0; k < 3; k++) { R[3 * b + k] = rx[k] * var; ang[3 * b + k] = ra[k]; }
        pos[3 * b] = rc[b].x; pos[3 * b + 1] = rc[b].y; pos[3 * b + 2] = rc[b].z;
    }
    const real dtg = (real)p->dt / p->gammaR;
    const real dta = (real)p->dt / ((real)p->gammaTheta * p->alpha);
    const real dtt = (real)p->dt / p->gammaTheta;
    const real vta = (real)p->varTheta * sqrt((real)p->freeze_temp / p->alpha);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t b = 0; b < nb; b++) {
        const int tr = traj_of(t, b);
        const int64_t base = t->off[tr], end = t->off[tr + 1];
        const real beta = beta_t[tr];
        real c[3], f[3];
        for (int k = 0; k < 3; k++) {
            c[k] = 1 / sqrt(1 + beta * beta * ci_bead[3 * b + k]);
            f[k] = F[7 * b + k] + R[3 * b + k] * c[k];
            c[k] *= beta;
        }
        const real L = t->box ? t->box[tr] : 0;
        for (int64_t j = base; j < end; j++) {
            if (j == b || t->extra[b] || t->extra[j]) continue;
            real e[3] = {pos[3 * j] - pos[3 * b], pos[3 * j + 1] - pos[3 * b + 1], pos[3 * j + 2] - pos[3 * b + 2]};
            min_image(e, L);
            const real r = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
            for (int k = 0; k < 3; k++) e[k] /= r;
            real D[9], v[3];
            rpy_hat(p->tea_a, e, r, D);
            for (int k = 0; k < 3; k++) v[k] = F[7 * j + k] + R[3 * j + k] * c[k];
            for (int x = 0; x < 3; x++) f[x] += D[3 * x] * v[0] + D[3 * x + 1] * v[1] + D[3 * x + 2] * v[2];
        }
        if (!t->fixed[b] && !t->extra[b]) {
            or_coord *q = &rc[b];
            q->x = (float)(pos[3 * b] + dtg * f[0]);
            q->y = (float)(pos[3 * b + 1] + dtg * f[1]);
            q->z = (float)(pos[3 * b + 2] + dtg * f[2]);
            const double *fa = F + 7 * b;
            q->fi = (float)((real)q->fi + dta * (real)fa[3] + vta * ang[3 * b]);
            q->psi = (float)((real)q->psi + dta * (real)fa[5] + vta * ang[3 * b + 1]);
            q->theta = (float)((real)q->theta + dtt * (real)fa[4] + (real)p->varTheta * ang[3 * b + 2]);
        }
    }
    free(R);
    free(ang);
    free(pos);
}

/* ------------------------------------------------------------- step loop */
/* The per-step order of compute(), src/compute_cuda.cu:1137-1240, without
 * the host events (stride output, hydrolysis, concentration control):
 *   if step % LJPairsUpdateFreq == 0: LJ list (lj_on), bonded lists (is_assembly)
 *   forces
 *   TEA: if step % tea_epsilon_freq == 0 refresh eps/beta; TEA step
 *   else: Euler-Maruyama step.
 * noise: nsteps * nb * 6 floats, or NULL to use `seeds`.  beta_t/ci_bead
 * carry TEA state across calls (may be NULL when TEA is off).  Returns 0, a
 * positive list-overflow count (first occurrence) or -(1 + traj) on a TEA
 * capricious abort. */
int64_t or_run(const or_params *p, const or_topo *t, or_coord *rc, int64_t nb, int64_t step0, int64_t nsteps,
               const float *noise, uint32_t *seeds, double *ci_bead, double *eps_t, double *beta_t)
{
    double *F = (double *)malloc(sizeof(double) * 7 * nb);
    double *epsb = p->tea_on ? (double *)malloc(sizeof(double) * nb) : NULL;
    int64_t ret = 0;
    for (int64_t s = 0; s < nsteps; s++) {
        const int64_t step = step0 + s;
        if (step % p->ljpairsupdatefreq == 0) {
            int64_t ov = 0;
            if (p->lj_on) ov += or_lj_list(p, t, rc, nb);
            if (p->is_assembly) ov += or_pairs(t, rc, nb);
            if (ov && !ret) ret = ov;
        }
        or_forces(p, t, rc, nb, F);
        const float *nz = noise ? noise + (size_t)6 * nb * s : NULL;
        if (p->tea_on) {
            if (step % p->tea_epsilon_freq == 0) {
                or_tea_eps(p, t, rc, nb, ci_bead, epsb);
                int e = or_tea_beta(p, t, epsb, eps_t, beta_t);
                if (e) { ret = -e; break; }
            }
            or_tea_integrate(p, t, rc, F, nb, ci_bead, beta_t, nz, seeds);
        } else {
            or_integrate(p, t, rc, F, nb, nz, seeds);
        }
    }
    free(F);
    free(epsb);
    return ret;
}

int or_real_size(void) { return (int)sizeof(real); }
int or_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
