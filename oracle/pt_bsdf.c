/*
 * pt_bsdf.c -- FP64 restatement of the reference's GGX conductor and smooth
 * dielectric scattering math.  TEST INFRASTRUCTURE ONLY (pt_bsdf.h lists
 * the reference lines followed).
 */
#include "pt_bsdf.h"

#include "ggx_lut_data.h"

#include <math.h>
#include <pthread.h>

#define NZ 1e-12
#define PI_D 3.14159265358979323846
#define INV_PI 0.31830988618379067154
#define TWO_PI 6.28318530717958647692
#define LUT RISE_GGX_LUT_SIZE

static const double ESS_FLAT[LUT * LUT] = {RISE_GGX_ESS_VALUES};
#define ESS_ROW(a) (ESS_FLAT + (a) * LUT)
static const double EAVG[LUT] = {RISE_GGX_EAVG_VALUES};
static const double GL_X[21] = {RISE_GL21_NODES};
static const double GL_W[21] = {RISE_GL21_WEIGHTS};

/* ---- vectors ---------------------------------------------------------- */

static rb_v3 mk(double x, double y, double z) { rb_v3 r = {x, y, z}; return r; }
static rb_v3 add(rb_v3 a, rb_v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static rb_v3 sub(rb_v3 a, rb_v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static rb_v3 mul(rb_v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static rb_v3 neg(rb_v3 a) { return mk(-a.x, -a.y, -a.z); }
static double dot(rb_v3 a, rb_v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static double mag(rb_v3 a) { return sqrt(dot(a, a)); }
static rb_v3 normalize(rb_v3 a)
{
    const double m = mag(a);
    if (m > 0.0) return mul(a, 1.0 / m);
    return a;
}
static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }
static rb_v3 to_local(const rb_onb* b, rb_v3 v) { return mk(dot(v, b->u), dot(v, b->v), dot(v, b->w)); }
static rb_v3 from_local(const rb_onb* b, rb_v3 a)
{
    return add(add(mul(b->u, a.x), mul(b->v, a.y)), mul(b->w, a.z));
}
/* OrthonormalBasis3D::Transform: the same sum, written componentwise. */
static rb_v3 onb_transform(const rb_onb* b, rb_v3 a)
{
    return mk(b->u.x * a.x + b->v.x * a.y + b->w.x * a.z, b->u.y * a.x + b->v.y * a.y + b->w.y * a.z,
              b->u.z * a.x + b->v.z * a.y + b->w.z * a.z);
}

/* ---- Smith / GGX (MicrofacetUtils.h:197-371) ------------------------------ */

static double lambda_aniso(double ax, double ay, rb_v3 v)
{
    const double vz2 = v.z * v.z;
    if (vz2 < 1e-20) return 1e10;
    ax = ax < 1e-4 ? 1e-4 : ax;
    ay = ay < 1e-4 ? 1e-4 : ay;
    const double a2 = ax * ax * v.x * v.x + ay * ay * v.y * v.y;
    return (-1.0 + sqrt(1.0 + a2 / vz2)) * 0.5;
}

static double g1(double ax, double ay, rb_v3 v) { return 1.0 / (1.0 + lambda_aniso(ax, ay, v)); }

static double g2(double ax, double ay, rb_v3 wi, rb_v3 wo)
{
    return 1.0 / (1.0 + lambda_aniso(ax, ay, wi) + lambda_aniso(ax, ay, wo));
}

static double ggx_d(double ax, double ay, rb_v3 h)
{
    if (h.z <= 0) return 0.0;
    ax = ax < 1e-4 ? 1e-4 : ax;
    ay = ay < 1e-4 ? 1e-4 : ay;
    const double hx = h.x / ax, hy = h.y / ay;
    const double denom = hx * hx + hy * hy + h.z * h.z;
    return 1.0 / (PI_D * ax * ay * denom * denom);
}

static rb_v3 vndf_sample(rb_v3 wi, const rb_onb* b, double ax, double ay, double u1, double u2)
{
    const double aeff = sqrt(ax * ay);
    if (aeff < 1e-6) return b->w;
    const rb_v3 wl = to_local(b, wi);
    const rb_v3 wh = normalize(mk(ax * wl.x, ay * wl.y, wl.z));
    const double phi = TWO_PI * u1;
    const double z = (1.0 - u2) * (1.0 + wh.z) - wh.z;
    const double st = sqrt(dmax(0.0, 1.0 - z * z));
    const double x = st * cos(phi), y = st * sin(phi);
    const rb_v3 c = mk(x + wh.x, y + wh.y, z + wh.z);
    const rb_v3 ml = normalize(mk(ax * c.x, ay * c.y, c.z));
    return normalize(from_local(b, ml));
}

static double vndf_pdf(rb_v3 wi, rb_v3 wo, const rb_onb* b, double ax, double ay)
{
    const double aeff = sqrt(ax * ay);
    if (aeff < 1e-6) return 0.0;
    const double cos_wi = dot(wi, b->w);
    if (cos_wi < 1e-10) return 0.0;
    const rb_v3 h = normalize(add(wi, wo));
    const double wih = dot(wi, h), woh = dot(wo, h);
    if (wih <= 0 || woh < 1e-10) return 0.0;
    const rb_v3 hl = to_local(b, h), wil = to_local(b, wi);
    if (hl.z <= 0) return 0.0;
    const double dv = g1(ax, ay, wil) * wih * ggx_d(ax, ay, hl) / cos_wi;
    return dv / (4.0 * woh);
}

/* ---- energy tables (MicrofacetEnergyLUT.h) -------------------------------- */

static void alpha_index(double alpha, int* i0, int* i1, double* f)
{
    const double a = dmax(0.0, dmin(1.0, (alpha - 0.01) / (1.0 - 0.01))) * (LUT - 1);
    *i0 = (int)a;
    *i1 = *i0 + 1 < LUT - 1 ? *i0 + 1 : LUT - 1;
    *f = a - *i0;
}

double rise_oracle_ggx_lookup_ess(double cos_theta, double alpha)
{
    int a0, a1;
    double af;
    alpha_index(alpha, &a0, &a1, &af);
    double c = dmax(0.0, dmin(1.0, cos_theta)) * LUT - 0.5;
    if (c < 0) c = 0;
    const int c0 = (int)c;
    const int c1 = c0 + 1 < LUT - 1 ? c0 + 1 : LUT - 1;
    const double cf = c - c0;
    const double v00 = ESS_ROW(a0)[c0], v01 = ESS_ROW(a0)[c1], v10 = ESS_ROW(a1)[c0], v11 = ESS_ROW(a1)[c1];
    return (1 - af) * ((1 - cf) * v00 + cf * v01) + af * ((1 - cf) * v10 + cf * v11);
}

double rise_oracle_ggx_lookup_eavg(double alpha)
{
    int a0, a1;
    double af;
    alpha_index(alpha, &a0, &a1, &af);
    return (1 - af) * EAVG[a0] + af * EAVG[a1];
}

/* Piecewise model of Ess(cos) for one alpha: a flat cap below the first
 * cell centre, linear spans between centres, a flat cap above the last. */
typedef struct { double lo, hi, ess_lo, slope; } seg;

static int build_segments(const double row[LUT], seg s[LUT + 1])
{
    int n = 0;
    const double c0 = 0.5 / LUT, cl = (LUT - 0.5) / LUT;
    s[n].lo = 0.0; s[n].hi = c0; s[n].ess_lo = row[0]; s[n].slope = 0.0; n++;
    for (int k = 0; k < LUT - 1; ++k) {
        const double ck = (k + 0.5) / LUT, ck1 = (k + 1.5) / LUT;
        s[n].lo = ck; s[n].hi = ck1; s[n].ess_lo = row[k]; s[n].slope = (row[k + 1] - row[k]) / (ck1 - ck); n++;
    }
    s[n].lo = cl; s[n].hi = 1.0; s[n].ess_lo = row[LUT - 1]; s[n].slope = 0.0; n++;
    return n;
}

static double seg_shape(const seg* s, double y) { return dmax(0.0, (1.0 - (s->ess_lo + s->slope * y)) * (s->lo + y)); }

/* Antiderivative of (1 - ess_lo - slope*y)(lo + y) from 0 to y. */
static double seg_cdf(const seg* s, double y)
{
    const double p = (1.0 - s->ess_lo) * s->lo;
    const double q = (1.0 - s->ess_lo) - s->slope * s->lo;
    const double r = -s->slope;
    return p * y + q * y * y * 0.5 + r * y * y * y * (1.0 / 3.0);
}

static double seg_invert(const seg* s, double target, double w)
{
    double ylo = 0.0, yhi = w, y = w * 0.5;
    for (int it = 0; it < 30; ++it) {
        const double fy = seg_cdf(s, y);
        const double diff = fy - target;
        if (fabs(diff) < 1e-13) break;
        if (diff > 0) yhi = y;
        else ylo = y;
        const double d = seg_shape(s, y);
        double yn = (d > 1e-12) ? (y - diff / d) : 0.5 * (ylo + yhi);
        if (!(yn > ylo && yn < yhi)) yn = 0.5 * (ylo + yhi);
        y = yn;
    }
    return y;
}

static double row_z[LUT];
static pthread_once_t row_z_once = PTHREAD_ONCE_INIT;

static void init_row_z(void)
{
    for (int r = 0; r < LUT; ++r) {
        seg s[LUT + 1];
        const int n = build_segments(ESS_ROW(r), s);
        double I = 0.0;
        for (int i = 0; i < n; ++i) I += seg_cdf(&s[i], s[i].hi - s[i].lo);
        row_z[r] = 2.0 * I;
    }
}

double rise_oracle_ggx_ms_z(double alpha)
{
    pthread_once(&row_z_once, init_row_z);
    int a0, a1;
    double af;
    alpha_index(alpha, &a0, &a1, &af);
    return (1.0 - af) * row_z[a0] + af * row_z[a1];
}

double rise_oracle_ggx_sample_ms_cos(double alpha, double u)
{
    int a0, a1;
    double af;
    alpha_index(alpha, &a0, &a1, &af);
    double row[LUT];
    for (int k = 0; k < LUT; ++k) row[k] = (1 - af) * ESS_ROW(a0)[k] + af * ESS_ROW(a1)[k];
    seg s[LUT + 1];
    const int n = build_segments(row, s);
    double tot[LUT + 1], I = 0.0;
    for (int i = 0; i < n; ++i) {
        tot[i] = seg_cdf(&s[i], s[i].hi - s[i].lo);
        I += tot[i];
    }
    if (I <= 1e-15) return dmax(0.0, dmin(1.0, u));
    const double target = dmax(0.0, dmin(I, u * I));
    int k = 0;
    double run = 0.0;
    while (k < n - 1 && run + tot[k] < target) {
        run += tot[k];
        k++;
    }
    const double lt = dmax(0.0, dmin(tot[k], target - run));
    const double y = seg_invert(&s[k], lt, s[k].hi - s[k].lo);
    return dmax(0.0, dmin(1.0, s[k].lo + y));
}

double rise_oracle_ggx_ms_pdf(double cos_theta, double alpha, double z)
{
    if (z <= 1e-15) return 0.0;
    const double c = dmax(0.0, dmin(1.0, cos_theta));
    return dmax(0.0, (1.0 - rise_oracle_ggx_lookup_ess(c, alpha)) * c / (PI_D * z));
}

/* ---- Fresnel (Optics.hpp:35-84, MicrofacetEnergyLUT.h:384-426) ----------- */

double rise_oracle_fresnel_conductor(double cos_i, double ni, double nt, double kt)
{
    double c = fabs(cos_i);
    if (c < NZ) c = NZ;
    const double c2 = c * c;
    double s2 = 1 - c2;
    if (s2 < 0.0) s2 = 0.0;
    const double s = sqrt(s2), t = s / c, t2 = t * t;
    const double ni2 = ni * ni, nt2 = nt * nt, kt2 = kt * kt;
    const double ns = ni2 * s2, nk = 4 * nt2 * kt2;
    double a2 = nt2 - kt2 - ns;
    const double sq = sqrt(a2 * a2 + nk);
    double b2 = sq - nt2 + kt2 + ns;
    a2 = sq + nt2 - kt2 - ns;
    const double dni = 0.5 / ni2;
    a2 *= dni;
    b2 *= dni;
    if (a2 < 0) a2 = 0;
    if (b2 < 0) b2 = 0;
    const double a = sqrt(a2), ab2 = a2 + b2, dcos = 2 * a * c, astan = 2 * a * s * t, stan = s2 * t2;
    const double rs = (ab2 - dcos + c2) / (ab2 + dcos + c2);
    const double rp = rs * (ab2 - astan + stan) / (ab2 + astan + stan);
    return 0.5 * (rp + rs);
}

double rise_oracle_fresnel_avg(double ni, double nt, double kt)
{
    /* Direction at angle acos(mu) from the normal: |v.n| = mu. */
    double sum = 0.0;
    for (int i = 0; i < 21; ++i) {
        const double mu = GL_X[i];
        sum += rise_oracle_fresnel_conductor(mu, ni, nt, kt) * (2.0 * mu * GL_W[i]);
    }
    return sum;
}

double rise_oracle_ggx_fms(double f_avg, double e_avg)
{
    const double denom = 1.0 - f_avg * (1.0 - e_avg);
    if (denom < 1e-10) return 0.0;
    return f_avg * f_avg * e_avg / denom;
}

/* ---- GGX material (GGXBRDF::valueNM, GGXSPF::PdfNM / ScatterNM) ----------- */

/* The ray-facing ONB (FlipW: w and u negated) and the geometric normal
 * oriented against the ray. */
static void facing(const rb_ggx* g, rb_onb* b, rb_v3* geom_n)
{
    *b = g->onb;
    if (dot(g->dir, g->onb.w) > NZ) {
        b->w = neg(b->w);
        b->u = neg(b->u);
    }
    const rb_v3 raw = (dot(g->ng, g->ng) > 1e-12) ? g->ng : b->w;
    *geom_n = (dot(raw, g->dir) < 0) ? raw : neg(raw);
}

double rb_ggx_value(const rb_ggx* g, rb_v3 vin)
{
    rb_onb b;
    rb_v3 gn;
    facing(g, &b, &gn);
    const rb_v3 n = b.w;
    const rb_v3 v = normalize(vin);
    const rb_v3 r = normalize(neg(g->dir));
    const double nr = dot(n, r), nv = dot(n, v);
    if (nr < NZ || nv < NZ) return 0.0;
    if (dot(v, gn) <= 0 || dot(r, gn) <= 0) return 0.0;
    const double ax = dmax(g->alpha, 1e-4), ay = ax;
    const rb_v3 h = normalize(add(v, r));
    const double D = ggx_d(ax, ay, to_local(&b, h));
    const double G2 = g2(ax, ay, to_local(&b, v), to_local(&b, r));
    const double spec_factor = D * G2 / (4.0 * nv * nr);
    double spec = 0.0;
    if (spec_factor > 0) {
        const double f = rise_oracle_fresnel_conductor(dot(g->dir, h), g->amb, g->ior, g->ext);
        if (f > 0) spec = g->rs * f * spec_factor;
    }
    const double aeff = sqrt(ax * ay);
    const double eavg = rise_oracle_ggx_lookup_eavg(aeff);
    if ((1.0 - eavg) > 1e-10) {
        const double ess_o = rise_oracle_ggx_lookup_ess(nr, aeff);
        const double ess_i = rise_oracle_ggx_lookup_ess(nv, aeff);
        const double f_ms_shape = (1.0 - ess_o) * (1.0 - ess_i) / (PI_D * (1.0 - eavg));
        const double favg = rise_oracle_fresnel_avg(g->amb, g->ior, g->ext);
        spec += rise_oracle_ggx_fms(g->rs * favg, eavg) * f_ms_shape;
    }
    return g->rd * INV_PI + spec;
}

double rb_ggx_pdf(const rb_ggx* g, rb_v3 wo)
{
    rb_onb b;
    rb_v3 gn;
    facing(g, &b, &gn);
    const rb_v3 n = b.w;
    const rb_v3 won = normalize(wo);
    const double ct = dot(won, n);
    if (ct <= 0) return 0.0;
    if (dot(won, gn) <= 0) return 0.0;
    const rb_v3 wi = normalize(neg(g->dir));
    const double ax = dmax(g->alpha, 1e-4), ay = ax, aeff = sqrt(ax * ay);
    const double wd = g->rd, ws = g->rs;
    const double wms = ws * (1.0 - rise_oracle_ggx_lookup_ess(dot(wi, n), aeff));
    const double total = wd + ws + wms;
    if (total < 1e-10) return ct * INV_PI;
    const double diff_pdf = ct * INV_PI;
    const double spec_pdf = aeff >= 1e-6 ? vndf_pdf(wi, won, &b, ax, ay) : 0.0;
    const double ms_pdf = rise_oracle_ggx_ms_pdf(ct, aeff, rise_oracle_ggx_ms_z(aeff));
    return (wd * diff_pdf + wms * ms_pdf + ws * spec_pdf) / total;
}

rb_sample rb_ggx_scatter(const rb_ggx* g, double u_lobe, double u1, double u2)
{
    rb_sample out = {0, {0, 0, 0}, 0, 0, 0};
    rb_onb b;
    rb_v3 gn;
    facing(g, &b, &gn);
    const rb_v3 n = b.w;
    const rb_v3 wi = normalize(neg(g->dir));
    const double ax = dmax(g->alpha, 1e-4), ay = ax, aeff = sqrt(ax * ay);
    const double wd = g->rd, ws = g->rs;
    const double eavg = rise_oracle_ggx_lookup_eavg(aeff);
    const double cos_wi = dot(wi, n);
    const double ess_i = rise_oracle_ggx_lookup_ess(cos_wi, aeff);
    const double wms = ws * (1.0 - ess_i);
    const double total = wd + ws + wms;
    const double msz = rise_oracle_ggx_ms_z(aeff);
    const double p_diff = (total > 1e-10) ? wd / total : 1.0;
    const double p_spec = (total > 1e-10) ? ws / total : 0.0;
    if (u_lobe < p_diff) {
        /* CreateDiffuseVector (GeometricUtilities.cpp:85-98). */
        const double phi = TWO_PI * u1, st = sqrt(u2), ctt = sqrt(1.0 - u2);
        const rb_v3 wo = onb_transform(&b, mk(cos(phi) * st, sin(phi) * st, ctt));
        const double ct = dot(wo, n);
        if (ct > 0 && dot(wo, gn) > 0) {
            const double diff_pdf = ct * INV_PI;
            const double spec_pdf = aeff >= 1e-6 ? vndf_pdf(wi, wo, &b, ax, ay) : 0.0;
            const double ms_pdf = rise_oracle_ggx_ms_pdf(ct, aeff, msz);
            out.ok = 1;
            out.dir = wo;
            out.pdf = total > 1e-10 ? (wd * diff_pdf + wms * ms_pdf + ws * spec_pdf) / total : diff_pdf;
            out.kray = wd / p_diff;
            out.type = 0;
        }
    } else if (u_lobe < p_diff + p_spec) {
        if (aeff >= 1e-6) {
            const rb_v3 m = vndf_sample(wi, &b, ax, ay, u1, u2);
            const double wim = dot(wi, m);
            if (wim > 0) {
                const rb_v3 wo = normalize(sub(mul(m, 2.0 * wim), wi));
                const double ct = dot(wo, n);
                if (ct > 0 && dot(wo, gn) > 0) {
                    const double vp = vndf_pdf(wi, wo, &b, ax, ay);
                    if (vp > 1e-10) {
                        const double diff_pdf = ct * INV_PI;
                        const double ms_pdf = rise_oracle_ggx_ms_pdf(ct, aeff, msz);
                        const double mix = total > 1e-10 ? (wd * diff_pdf + wms * ms_pdf + ws * vp) / total : vp;
                        const double F = g->rs * rise_oracle_fresnel_conductor(dot(g->dir, m), g->amb, g->ior, g->ext);
                        const double G2 = g2(ax, ay, to_local(&b, wi), to_local(&b, wo));
                        const double G1 = g1(ax, ay, to_local(&b, wi));
                        const double kray = (G1 > 1e-10) ? F * G2 / (G1 * p_spec) : 0.0;
                        if (kray > 0) {
                            out.ok = 1;
                            out.dir = wo;
                            out.pdf = mix;
                            out.kray = kray;
                            out.type = 1;
                        }
                    }
                }
            }
        }
    } else {
        const double p_ms = 1.0 - p_diff - p_spec;
        if (p_ms > 1e-10 && (1.0 - eavg) > 1e-10) {
            const double ct = rise_oracle_ggx_sample_ms_cos(aeff, u1);
            const double st = sqrt(dmax(0.0, 1.0 - ct * ct));
            const double phi = TWO_PI * u2;
            const rb_v3 wo = onb_transform(&b, mk(st * cos(phi), st * sin(phi), ct));
            if (ct > 0 && dot(wo, gn) > 0) {
                const double diff_pdf = ct * INV_PI;
                const double spec_pdf = aeff >= 1e-6 ? vndf_pdf(wi, wo, &b, ax, ay) : 0.0;
                const double ms_pdf = rise_oracle_ggx_ms_pdf(ct, aeff, msz);
                const double mix = total > 1e-10 ? (wd * diff_pdf + wms * ms_pdf + ws * spec_pdf) / total : diff_pdf;
                const double ess_o = rise_oracle_ggx_lookup_ess(ct, aeff);
                const double favg = rise_oracle_fresnel_avg(g->amb, g->ior, g->ext);
                const double fms = rise_oracle_ggx_fms(g->rs * favg, eavg);
                const double f_ms = fms * (1.0 - ess_o) * (1.0 - ess_i) * INV_PI / (1.0 - eavg);
                const double kray = (ms_pdf > 1e-14) ? f_ms * ct / (ms_pdf * p_ms) : 0.0;
                if (kray > 0) {
                    out.ok = 1;
                    out.dir = wo;
                    out.pdf = mix;
                    out.kray = kray;
                    out.type = 0;
                }
            }
        }
    }
    return out;
}

/* ---- flat wrappers for tests ------------------------------------------- */

static rb_ggx ctx_of(const double c[21])
{
    rb_ggx g;
    g.dir = mk(c[0], c[1], c[2]);
    g.onb.u = mk(c[3], c[4], c[5]);
    g.onb.v = mk(c[6], c[7], c[8]);
    g.onb.w = mk(c[9], c[10], c[11]);
    g.ng = mk(c[12], c[13], c[14]);
    g.rd = c[15];
    g.rs = c[16];
    g.alpha = c[17];
    g.ior = c[18];
    g.ext = c[19];
    g.amb = c[20];
    return g;
}

double rise_oracle_ggx_value(const double ctx[21], const double vin[3])
{
    const rb_ggx g = ctx_of(ctx);
    return rb_ggx_value(&g, mk(vin[0], vin[1], vin[2]));
}

double rise_oracle_ggx_pdf(const double ctx[21], const double wo[3])
{
    const rb_ggx g = ctx_of(ctx);
    return rb_ggx_pdf(&g, mk(wo[0], wo[1], wo[2]));
}

void rise_oracle_ggx_scatter(const double ctx[21], double u_lobe, double u1, double u2, double out[7])
{
    const rb_ggx g = ctx_of(ctx);
    const rb_sample s = rb_ggx_scatter(&g, u_lobe, u1, u2);
    out[0] = s.ok;
    out[1] = s.dir.x;
    out[2] = s.dir.y;
    out[3] = s.dir.z;
    out[4] = s.kray;
    out[5] = s.pdf;
    out[6] = s.type;
}

double rise_oracle_ggx_d(double ax, double ay, const double h[3]) { return ggx_d(ax, ay, mk(h[0], h[1], h[2])); }
double rise_oracle_ggx_g1(double ax, double ay, const double v[3]) { return g1(ax, ay, mk(v[0], v[1], v[2])); }
double rise_oracle_ggx_g2(double ax, double ay, const double wi[3], const double wo[3])
{
    return g2(ax, ay, mk(wi[0], wi[1], wi[2]), mk(wo[0], wo[1], wo[2]));
}

static rb_onb onb_of(const double o[9])
{
    rb_onb b;
    b.u = mk(o[0], o[1], o[2]);
    b.v = mk(o[3], o[4], o[5]);
    b.w = mk(o[6], o[7], o[8]);
    return b;
}

void rise_oracle_vndf_sample(const double wi[3], const double onb9[9], double ax, double ay, double u1, double u2,
                             double m_out[3])
{
    const rb_onb b = onb_of(onb9);
    const rb_v3 m = vndf_sample(mk(wi[0], wi[1], wi[2]), &b, ax, ay, u1, u2);
    m_out[0] = m.x;
    m_out[1] = m.y;
    m_out[2] = m.z;
}

double rise_oracle_vndf_pdf(const double wi[3], const double wo[3], const double onb9[9], double ax, double ay)
{
    const rb_onb b = onb_of(onb9);
    return vndf_pdf(mk(wi[0], wi[1], wi[2]), mk(wo[0], wo[1], wo[2]), &b, ax, ay);
}

/* ---- optics (Optics.cpp:20-160) ------------------------------------------ */

rb_v3 rb_reflect(rb_v3 vin, rb_v3 n)
{
    const double m = mag(n);
    if (m < NZ) return vin;
    if (fabs(m - 1.0) > 1e-6) n = normalize(n);
    const double d = 2.0 * dot(n, vin);
    return mk(vin.x - d * n.x, vin.y - d * n.y, vin.z - d * n.z);
}

int rb_refract(rb_v3 n, double ni, double nt, rb_v3* vin)
{
    if (ni <= NZ || nt <= NZ) return 0;
    const double nm = mag(n), im = mag(*vin);
    if (nm < NZ || im < NZ) return 0;
    rb_v3 un = n, ui = *vin;
    if (fabs(nm - 1.0) > 1e-6) un = normalize(un);
    if (fabs(im - 1.0) > 1e-6) ui = normalize(ui);
    if (dot(un, ui) > 0) un = neg(un);
    double k = dot(un, ui);
    const rb_v3 s = mul(sub(ui, mul(un, k)), ni / nt);
    k = 1.0 - dot(s, s);
    if (k < NZ) return 0;
    *vin = sub(s, mul(un, sqrt(k)));
    return 1;
}

double rb_dielectric_reflectance(rb_v3 v, rb_v3 tv, rb_v3 n, double ni, double nt)
{
    if (ni <= NZ || nt <= NZ) return 1.0;
    const double nm = mag(n), vm = mag(v), tm = mag(tv);
    if (nm < NZ || vm < NZ || tm < NZ) return 1.0;
    if (fabs(nm - 1.0) > 1e-6) n = normalize(n);
    if (fabs(vm - 1.0) > 1e-6) v = normalize(v);
    if (fabs(tm - 1.0) > 1e-6) tv = normalize(tv);
    const double ci = fabs(dot(v, n)), ct = fabs(dot(tv, n));
    const double nn = ni * nt, cc = ct * ci, ni2 = ni * ni, nt2 = nt * nt, ci2 = ci * ci, ct2 = ct * ct;
    const double nit = ni2 - nt2, cit = ci2 - ct2;
    const double num = nit * nit * cc * cc + cit * cit * nn * nn;
    double den = cc * (ni2 + nt2) + nn * (ci2 + ct2);
    den = den * den;
    double r;
    if ((den < 0.000001) && (num < 0.000001)) r = 1.0;
    else r = num / den;
    return r < 1.0 ? r : 1.0;
}

double rise_oracle_dielectric_reflectance(const double v[3], const double tv[3], const double n[3], double ni,
                                          double nt)
{
    return rb_dielectric_reflectance(mk(v[0], v[1], v[2]), mk(tv[0], tv[1], tv[2]), mk(n[0], n[1], n[2]), ni, nt);
}

int rise_oracle_refract(const double n[3], double ni, double nt, double vin[3])
{
    rb_v3 v = mk(vin[0], vin[1], vin[2]);
    const int ok = rb_refract(mk(n[0], n[1], n[2]), ni, nt, &v);
    if (ok) {
        vin[0] = v.x;
        vin[1] = v.y;
        vin[2] = v.z;
    }
    return ok;
}

void rise_oracle_reflect(const double vin[3], const double n[3], double out[3])
{
    const rb_v3 r = rb_reflect(mk(vin[0], vin[1], vin[2]), mk(n[0], n[1], n[2]));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}
