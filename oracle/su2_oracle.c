/*
 * su2_oracle.c -- CPU oracle for the pure-gauge HMC molecular-dynamics path of
 * HISKP-LQCD/su2.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this library, and only as the checker.  The
 * product library (su2_amd/libsu2mi355x.so) never links or calls it.
 *
 * A plain-C restatement of the reference's algorithm.  Every routine cites the
 * reference file:line it follows.  Arithmetic is written in the reference's own
 * operation order (std::complex products as (ac-bd, ad+bc), left-to-right
 * SU(2) products, loop orders of the OpenMP loops) and the library is built
 * with -ffp-contract=off, so on x86-64 it reproduces the reference's bits; the
 * reference build in oracle/_ref is used by the tests to show that.
 *
 * Storage (reference AoS, gaugeconfig.hh:211-222): link (t,x,y,z,mu) lives at
 * ((((t*Lx+x)*Ly+y)*Lz+z)*nd+mu).  An SU(2) link is 4 doubles
 * (Re a, Im a, Re b, Im b) for [[a,b],[-b*,a*]] (su2.hh:11-70); a U(1) link is
 * one phase (u1.hh:11-60).  Momenta are 3 doubles (a,b,c) per SU(2) link
 * (adjointfield.hh:12-51) or one double per U(1) link.
 *
 * Random streams: std::mt19937 ([rand.eng.mers], ISO C++), libstdc++'s
 * generate_canonical<double,53>, uniform_real_distribution and the Marsaglia
 * polar normal_distribution (GCC 11 libstdc++ bits/random.tcc).  Pinned by the
 * standard's known answer (10000th output 4123659995) and by the reference
 * build's draws.
 *
 * dims[] = {Lx, Ly, Lz, Lt, nd} everywhere (gaugeconfig.hh:49-63 argument
 * order); g = 0 for U(1), 1 for SU(2).
 */
#define _GNU_SOURCE /* sincos */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* random numbers                                                           */
/* ------------------------------------------------------------------------ */

typedef struct {
  uint32_t mt[624];
  int idx;
} orc_mt19937;

typedef struct {
  int saved_available;
  double saved;
} orc_normal_dist;

ORC_API size_t orc_sizeof_rng(void) { return sizeof(orc_mt19937); }

/* std::mt19937::seed: value taken mod 2^32 (uint_fast32_t is 64 bit here). */
ORC_API void orc_mt_seed(orc_mt19937 *e, uint64_t seed) {
  e->mt[0] = (uint32_t)(seed & 0xffffffffu);
  for (int i = 1; i < 624; ++i)
    e->mt[i] = 1812433253u * (e->mt[i - 1] ^ (e->mt[i - 1] >> 30)) + (uint32_t)i;
  e->idx = 624;
}

static void mt_twist(orc_mt19937 *e) {
  const uint32_t upper = 0x80000000u, lower = 0x7fffffffu, mat = 0x9908b0dfu;
  for (int k = 0; k < 624; ++k) {
    uint32_t y = (e->mt[k] & upper) | (e->mt[(k + 1) % 624] & lower);
    e->mt[k] = e->mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? mat : 0u);
  }
  e->idx = 0;
}

ORC_API uint32_t orc_mt_next(orc_mt19937 *e) {
  if (e->idx >= 624) mt_twist(e);
  uint32_t y = e->mt[e->idx++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

/* generate_canonical<double, 53>(mt19937): two 32-bit words, (w0 + w1*2^32)/2^64,
 * clamped below 1 (libstdc++ bits/random.tcc). */
ORC_API double orc_canonical(orc_mt19937 *e) {
  double sum = 0.0, tmp = 1.0;
  for (int k = 0; k < 2; ++k) {
    sum += (double)orc_mt_next(e) * tmp;
    tmp *= 4294967296.0;
  }
  double r = sum / tmp;
  if (r >= 1.0) r = nextafter(1.0, 0.0);
  return r;
}

/* uniform_real_distribution<double>(a,b): canonical*(b-a)+a. */
static double uniform(orc_mt19937 *e, double a, double b) {
  return orc_canonical(e) * (b - a) + a;
}

/* normal_distribution<double>(0,1), Marsaglia polar method as in libstdc++:
 * returns y*mult first and caches x*mult. */
ORC_API double orc_normal(orc_mt19937 *e, orc_normal_dist *d) {
  double ret;
  if (d->saved_available) {
    d->saved_available = 0;
    ret = d->saved;
  } else {
    double x, y, r2;
    do {
      x = 2.0 * orc_canonical(e) - 1.0;
      y = 2.0 * orc_canonical(e) - 1.0;
      r2 = x * x + y * y;
    } while (r2 > 1.0 || r2 == 0.0);
    const double mult = sqrt(-2 * log(r2) / r2);
    d->saved = x * mult;
    d->saved_available = 1;
    ret = y * mult;
  }
  return ret * 1.0 + 0.0;
}

/* first n normals of one normal_distribution over mt19937(seed) */
ORC_API void orc_normals(uint64_t seed, size_t n, double *out) {
  orc_mt19937 e;
  orc_normal_dist d = {0, 0.0};
  orc_mt_seed(&e, seed);
  for (size_t i = 0; i < n; ++i) out[i] = orc_normal(&e, &d);
}

ORC_API void orc_mt_words(uint64_t seed, size_t n, uint32_t *out) {
  orc_mt19937 e;
  orc_mt_seed(&e, seed);
  for (size_t i = 0; i < n; ++i) out[i] = orc_mt_next(&e);
}

/* pi() of random_element.hh:7: atan(1)*4, the double nearest pi */
static const double ORC_PI = 3.14159265358979311599796346854;

/* ------------------------------------------------------------------------ */
/* SU(2) and U(1) algebra (su2.hh:11-123, u1.hh:11-119)                     */
/* ------------------------------------------------------------------------ */

typedef struct { double ar, ai, br, bi; } su2;
typedef struct { double re, im; } cplx;

static inline cplx cmul(cplx x, cplx y) {
  cplx r = {x.re * y.re - x.im * y.im, x.re * y.im + x.im * y.re};
  return r;
}
static inline cplx cadd(cplx x, cplx y) { cplx r = {x.re + y.re, x.im + y.im}; return r; }
static inline cplx csub(cplx x, cplx y) { cplx r = {x.re - y.re, x.im - y.im}; return r; }
static inline cplx cconj(cplx x) { cplx r = {x.re, -x.im}; return r; }
/* std::exp(th * Complex(0,1)) (u1.hh:27-29,95-119): glibc cexp of (th*0, th).
 * cexp is called rather than cos/sin because libm's public sin/cos may be a
 * different (FMA) variant than the one cexp uses internally. */
static inline cplx cexpi(double th) {
  double complex z = cexp(CMPLX(th * 0.0, th));
  cplx r = {creal(z), cimag(z)};
  return r;
}

/* operator*(su2,su2): a = a1 a2 - b1 conj(b2), b = a1 b2 + b1 conj(a2)  (su2.hh:87-92) */
static inline su2 su2_mul(su2 u, su2 v) {
  cplx a1 = {u.ar, u.ai}, b1 = {u.br, u.bi}, a2 = {v.ar, v.ai}, b2 = {v.br, v.bi};
  cplx a = csub(cmul(a1, a2), cmul(b1, cconj(b2)));
  cplx b = cadd(cmul(a1, b2), cmul(b1, cconj(a2)));
  su2 r = {a.re, a.im, b.re, b.im};
  return r;
}
/* dagger: (conj a, -b)  (su2.hh:53-55) */
static inline su2 su2_dag(su2 u) { su2 r = {u.ar, -u.ai, -u.br, -u.bi}; return r; }
static inline su2 su2_add(su2 u, su2 v) {
  su2 r = {u.ar + v.ar, u.ai + v.ai, u.br + v.br, u.bi + v.bi};
  return r;
}
/* Complex(f,0) * su2 (su2.hh:108-113): componentwise scale for finite values */
static inline su2 su2_scale(double f, su2 u) {
  su2 r = {u.ar * f, u.ai * f, u.br * f, u.bi * f};
  return r;
}
/* restoreSU (su2.hh:62-66): r = sqrt(|a|^2+|b|^2); a/=r; b/=r */
static inline su2 su2_restore(su2 u) {
  double na = hypot(u.ar, u.ai), nb = hypot(u.br, u.bi);
  double r = sqrt(na * na + nb * nb);
  /* complex<double> /= double divides both parts */
  su2 o = {u.ar / r, u.ai / r, u.br / r, u.bi / r};
  return o;
}
/* su2::round(n) (su2.hh:29-33) */
static inline su2 su2_round(su2 u, double dn) {
  su2 r = {round(u.ar * dn) / dn, round(u.ai * dn) / dn, round(u.br * dn) / dn,
           round(u.bi * dn) / dn};
  return r;
}
/* exp(adjointsu2) (exp_gauge.cc:8-18): (cos al + i s n2, s n1 + i s n0) */
static inline su2 su2_exp(double a, double b, double c) {
  const double alpha = sqrt(a * a + b * b + c * c);
  const double n0 = a / alpha, n1 = b / alpha, n2 = c / alpha;
  const double s = sin(alpha);
  su2 r = {cos(alpha), s * n2, s * n1, s * n0};
  return r;
}

static inline su2 ld_su2(const double *U, size_t i) {
  su2 r = {U[4 * i], U[4 * i + 1], U[4 * i + 2], U[4 * i + 3]};
  return r;
}
static inline void st_su2(double *U, size_t i, su2 u) {
  U[4 * i] = u.ar; U[4 * i + 1] = u.ai; U[4 * i + 2] = u.br; U[4 * i + 3] = u.bi;
}

/* ------------------------------------------------------------------------ */
/* geometry (gaugeconfig.hh:211-222)                                        */
/* ------------------------------------------------------------------------ */

typedef struct {
  long L[4]; /* extents indexed by coordinate: L[0]=Lt, L[1]=Lx, L[2]=Ly, L[3]=Lz */
  int nd;
  size_t V;
} geom;

static geom mkgeom(const int *dims) {
  geom g;
  g.L[1] = dims[0]; g.L[2] = dims[1]; g.L[3] = dims[2]; g.L[0] = dims[3];
  g.nd = dims[4];
  g.V = (size_t)g.L[0] * g.L[1] * g.L[2] * g.L[3];
  return g;
}

static inline size_t site_of(const geom *g, const long c[4]) {
  long y[4];
  for (int k = 0; k < 4; ++k) y[k] = ((c[k] % g->L[k]) + g->L[k]) % g->L[k];
  return ((size_t)((y[0] * g->L[1] + y[1]) * g->L[2] + y[2])) * g->L[3] + y[3];
}
static inline void coords_of(const geom *g, size_t s, long c[4]) {
  c[3] = s % g->L[3]; s /= g->L[3];
  c[2] = s % g->L[2]; s /= g->L[2];
  c[1] = s % g->L[1]; s /= g->L[1];
  c[0] = (long)s;
}
/* link index of (site shifted by sh[], mu) */
static inline size_t lk(const geom *g, size_t site, int mu) { return site * g->nd + mu; }
static inline size_t nb(const geom *g, const long c[4], int d1, int s1, int d2, int s2) {
  long y[4] = {c[0], c[1], c[2], c[3]};
  if (d1 >= 0) y[d1] += s1;
  if (d2 >= 0) y[d2] += s2;
  return site_of(g, y);
}

/* ------------------------------------------------------------------------ */
/* start configurations                                                     */
/* ------------------------------------------------------------------------ */

/* random_element (random_element.hh:9-25, SU(2); :27-34, U(1)) */
static su2 random_su2(orc_mt19937 *e, double delta) {
  const double alpha = uniform(e, 0.0, delta * 2 * ORC_PI);
  const double u = uniform(e, -1.0, 1.0);
  const double theta = uniform(e, 0.0, 2 * ORC_PI);
  const double r = sqrt(1 - u * u);
  /* one sincos per angle, as the reference's GCC build evaluates the pairs */
  double salpha, calpha, st, ct;
  sincos(alpha, &salpha, &calpha);
  sincos(theta, &st, &ct);
  const double sr = salpha * r;
  su2 o = {calpha, salpha * u, sr * st, sr * ct};
  return o;
}
static double random_u1(orc_mt19937 *e, double delta) {
  return uniform(e, -ORC_PI * delta, ORC_PI * delta);
}

/* hotstart (gaugeconfig.hh:277-302): delta clamped to [0,1], one random_element
 * per link in storage order from mt19937(seed). */
ORC_API void orc_hotstart(int grp, const int *dims, double *U, uint64_t seed, double delta) {
  geom g = mkgeom(dims);
  if (delta < 0.) delta = 0.;
  if (delta > 1.) delta = 1.;
  orc_mt19937 e;
  orc_mt_seed(&e, seed);
  size_t n = g.V * g.nd;
  for (size_t i = 0; i < n; ++i) {
    if (grp == 1) st_su2(U, i, random_su2(&e, delta));
    else U[i] = random_u1(&e, delta);
  }
}

/* coldstart (gaugeconfig.hh:254-266) */
ORC_API void orc_coldstart(int grp, const int *dims, double *U) {
  geom g = mkgeom(dims);
  size_t n = g.V * g.nd;
  for (size_t i = 0; i < n; ++i) {
    if (grp == 1) { su2 one = {1, 0, 0, 0}; st_su2(U, i, one); }
    else U[i] = 0.0;
  }
}

/* random_gauge_trafo (random_gauge_trafo.hh:10-33): sites in t,x,y,z order;
 * g(x) from random_element(.,1); U_mu(x) <- g U_mu(x); U_mu(x-mu) <- U_mu(x-mu) g^+ */
ORC_API void orc_random_gauge_trafo(int grp, const int *dims, double *U, uint64_t seed) {
  geom g = mkgeom(dims);
  orc_mt19937 e;
  orc_mt_seed(&e, seed);
  for (size_t s = 0; s < g.V; ++s) {
    long c[4];
    coords_of(&g, s, c);
    if (grp == 1) {
      su2 r = random_su2(&e, 1.0);
      for (int mu = 0; mu < g.nd; ++mu) {
        size_t i = lk(&g, s, mu);
        st_su2(U, i, su2_mul(r, ld_su2(U, i)));
        size_t j = lk(&g, nb(&g, c, mu, -1, -1, 0), mu);
        st_su2(U, j, su2_mul(ld_su2(U, j), su2_dag(r)));
      }
    } else {
      double r = random_u1(&e, 1.0);
      for (int mu = 0; mu < g.nd; ++mu) {
        size_t i = lk(&g, s, mu);
        U[i] = r + U[i];
        size_t j = lk(&g, nb(&g, c, mu, -1, -1, 0), mu);
        U[j] = U[j] + (-r);
      }
    }
  }
}

/* Haar-uniform SU(2) start (not in the reference): each link is a normalised
 * 4-vector of normal_distribution draws (Re a, Im a, Re b, Im b) from
 * mt19937(seed), links in storage order.  U(1): uniform phase in [-pi,pi).
 * Pinned by the reference build's normal stream (the first links are its
 * normalised 4-tuples) and by the Haar measure's moments, component law and
 * left invariance (tests/test_oracle.py). */
ORC_API void orc_haarstart(int grp, const int *dims, double *U, uint64_t seed) {
  geom g = mkgeom(dims);
  orc_mt19937 e;
  orc_normal_dist d = {0, 0.0};
  orc_mt_seed(&e, seed);
  size_t n = g.V * g.nd;
  for (size_t i = 0; i < n; ++i) {
    if (grp == 1) {
      double q[4], r2 = 0.0;
      for (int k = 0; k < 4; ++k) { q[k] = orc_normal(&e, &d); r2 += q[k] * q[k]; }
      double r = sqrt(r2);
      su2 o = {q[0] / r, q[1] / r, q[2] / r, q[3] / r};
      st_su2(U, i, o);
    } else {
      U[i] = uniform(&e, -ORC_PI, ORC_PI);
    }
  }
}

/* initnormal (adjointfield.hh:240-258): getSize()=V*nd links, components a,b,c
 * (SU(2)) or a (U(1)) from one normal_distribution in storage order. */
static void initnormal(int grp, const geom *g, double *P, orc_mt19937 *e) {
  orc_normal_dist d = {0, 0.0};
  size_t n = g->V * g->nd * (grp == 1 ? 3 : 1);
  for (size_t i = 0; i < n; ++i) P[i] = orc_normal(e, &d);
}
ORC_API void orc_initnormal(int grp, const int *dims, double *P, uint64_t seed) {
  geom g = mkgeom(dims);
  orc_mt19937 e;
  orc_mt_seed(&e, seed);
  initnormal(grp, &g, P, &e);
}

/* ------------------------------------------------------------------------ */
/* staples and the Wilson gauge force                                       */
/* ------------------------------------------------------------------------ */

/* get_staples (get_staples.hh:26-80) + gaugemonomial::derivative
 * (flat-gaugemonomial.hh:105-131) + get_deriv (adjointfield.hh:53-56,79-81):
 * deriv(x,mu) += (fac*beta/Nc) * get_deriv(U_mu(x) K_mu(x)). */
static void derivative_link(int grp, const geom *g, const double *U, size_t s, int mu,
                            double xi, int aniso, double coef, double *deriv) {
  long c[4];
  coords_of(g, s, c);
  const int nd = g->nd;
  if (grp == 1) {
    su2 K = {0, 0, 0, 0};
    for (int nu = 0; nu < nd; ++nu) {
      if (nu == mu) continue;
      double f = ((nu == 0) || (mu == 0)) ? 1.0 / xi : xi;
      su2 u1 = ld_su2(U, lk(g, nb(g, c, mu, 1, -1, 0), nu));
      if (aniso) u1 = su2_scale(f, u1);
      su2 t = su2_mul(su2_mul(u1, su2_dag(ld_su2(U, lk(g, nb(g, c, nu, 1, -1, 0), mu)))),
                      su2_dag(ld_su2(U, lk(g, s, nu))));
      K = su2_add(K, t);
    }
    for (int nu = 0; nu < nd; ++nu) {
      if (nu == mu) continue;
      double f = ((nu == 0) || (mu == 0)) ? 1.0 / xi : xi;
      su2 u1 = su2_dag(ld_su2(U, lk(g, nb(g, c, mu, 1, nu, -1), nu)));
      if (aniso) u1 = su2_scale(f, u1);
      size_t xm = nb(g, c, nu, -1, -1, 0);
      su2 t = su2_mul(su2_mul(u1, su2_dag(ld_su2(U, lk(g, xm, mu)))), ld_su2(U, lk(g, xm, nu)));
      K = su2_add(K, t);
    }
    su2 S = su2_mul(ld_su2(U, lk(g, s, mu)), K);
    double *d = deriv + 3 * lk(g, s, mu);
    d[0] += coef * (2. * S.bi);
    d[1] += coef * (2. * S.br);
    d[2] += coef * (2. * S.ai);
  } else {
    cplx K = {0, 0};
    for (int nu = 0; nu < nd; ++nu) {
      if (nu == mu) continue;
      double th1 = U[lk(g, nb(g, c, mu, 1, -1, 0), nu)];
      double th2 = U[lk(g, nb(g, c, nu, 1, -1, 0), mu)];
      double th3 = U[lk(g, s, nu)];
      if (!aniso) {
        K = cadd(K, cexpi((th1 + (-th2)) + (-th3)));
      } else {
        double f = ((nu == 0) || (mu == 0)) ? 1.0 / xi : xi;
        cplx fc = {f, 0.0};
        cplx t = cmul(cmul(cmul(fc, cexpi(th1)), cexpi(-th2)), cexpi(-th3));
        K = cadd(K, t);
      }
    }
    for (int nu = 0; nu < nd; ++nu) {
      if (nu == mu) continue;
      size_t xm = nb(g, c, nu, -1, -1, 0);
      double th1 = U[lk(g, nb(g, c, mu, 1, nu, -1), nu)];
      double th2 = U[lk(g, xm, mu)];
      double th3 = U[lk(g, xm, nu)];
      if (!aniso) {
        K = cadd(K, cexpi(((-th1) + (-th2)) + th3));
      } else {
        double f = ((nu == 0) || (mu == 0)) ? 1.0 / xi : xi;
        cplx fc = {f, 0.0};
        cplx t = cmul(cmul(cmul(fc, cexpi(-th1)), cexpi(-th2)), cexpi(th3));
        K = cadd(K, t);
      }
    }
    cplx S = cmul(cexpi(U[lk(g, s, mu)]), K);
    deriv[lk(g, s, mu)] += coef * S.im;
  }
}

ORC_API void orc_derivative(int grp, const int *dims, const double *U, double beta, double xi,
                            double fac, double *deriv) {
  geom g = mkgeom(dims);
  const int aniso = (xi != 1.0);
  const double Nc = (grp == 1) ? 2.0 : 1.0;
  const double coef = fac * beta / Nc;
#pragma omp parallel for schedule(static)
  for (size_t s = 0; s < g.V; ++s)
    for (int mu = 0; mu < g.nd; ++mu) derivative_link(grp, &g, U, s, mu, xi, aniso, coef, deriv);
}

/* ------------------------------------------------------------------------ */
/* MD updates                                                               */
/* ------------------------------------------------------------------------ */

static size_t nmom(int grp, const geom *g) { return g->V * g->nd * (grp == 1 ? 3 : 1); }

/* update_momenta (update_momenta.hh:10-33) with the gauge monomial as the only
 * md-active timescale-0 monomial; round_n>0 gives round_and_update_momenta
 * (:35-59). deriv is caller scratch of the momentum size. */
static void update_momenta(int grp, const geom *g, const double *U, double beta, double xi,
                           double *P, double *deriv, double dtau, double round_n) {
  size_t n = nmom(grp, g);
  memset(deriv, 0, n * sizeof(double));
  int dims[5] = {(int)g->L[1], (int)g->L[2], (int)g->L[3], (int)g->L[0], g->nd};
  orc_derivative(grp, dims, U, beta, xi, 1.0, deriv);
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) {
    double d = deriv[i];
    if (round_n > 0) d = round(d * round_n) / round_n;
    P[i] -= dtau * d;
  }
}

ORC_API void orc_update_momenta(int grp, const int *dims, const double *U, double beta,
                                double xi, double *P, double dtau) {
  geom g = mkgeom(dims);
  double *deriv = (double *)malloc(nmom(grp, &g) * sizeof(double));
  update_momenta(grp, &g, U, beta, xi, P, deriv, dtau, 0.0);
  free(deriv);
}

/* update_gauge (update_gauge.hh:8-16): U <- exp(dtau P) U; round_n>0 gives
 * round_and_update_gauge (:18-27) which rounds U before the product. */
static void update_gauge(int grp, const geom *g, double *U, const double *P, double dtau,
                         double round_n) {
  size_t n = g->V * g->nd;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) {
    if (grp == 1) {
      su2 u = ld_su2(U, i);
      if (round_n > 0) u = su2_round(u, round_n);
      st_su2(U, i, su2_mul(su2_exp(dtau * P[3 * i], dtau * P[3 * i + 1], dtau * P[3 * i + 2]), u));
    } else {
      double u = U[i];
      /* _u1::round (u1.hh:29-32) rounds the phase itself before scaling */
      if (round_n > 0) u = (double)(round(u) * round_n) / round_n;
      U[i] = dtau * P[i] + u;
    }
  }
}

ORC_API void orc_update_gauge(int grp, const int *dims, double *U, const double *P,
                              double dtau) {
  geom g = mkgeom(dims);
  update_gauge(grp, &g, U, P, dtau, 0.0);
}

/* gaugeconfig::restoreSU (gaugeconfig.hh:115-120); no-op for U(1) (u1.hh:54-55) */
ORC_API void orc_restoreSU(int grp, const int *dims, double *U) {
  geom g = mkgeom(dims);
  if (grp != 1) return;
  size_t n = g.V * g.nd;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) st_su2(U, i, su2_restore(ld_su2(U, i)));
}

/* ------------------------------------------------------------------------ */
/* sums                                                                     */
/* ------------------------------------------------------------------------ */

/* Neumaier-compensated accumulator (used only when compensated != 0). */
typedef struct { double s, c; } ksum;
static inline void kadd(ksum *k, double v, int comp) {
  if (!comp) { k->s += v; return; }
  double t = k->s + v;
  if (fabs(k->s) >= fabs(v)) k->c += (k->s - t) + v;
  else k->c += (v - t) + k->s;
  k->s = t;
}
static inline double kval(const ksum *k) { return k->s + k->c; }

static double plaq_retrace(int grp, const geom *g, const double *U, const long c[4], size_t s,
                           int mu, int nu) {
  size_t xpm = nb(g, c, mu, 1, -1, 0), xpn = nb(g, c, nu, 1, -1, 0);
  if (grp == 1) {
    su2 p = su2_mul(su2_mul(su2_mul(ld_su2(U, lk(g, s, mu)), ld_su2(U, lk(g, xpm, nu))),
                            su2_dag(ld_su2(U, lk(g, xpn, mu)))),
                    su2_dag(ld_su2(U, lk(g, s, nu))));
    return 2. * p.ar; /* retrace (su2.hh:72-75) */
  } else {
    double ph = ((U[lk(g, s, mu)] + U[lk(g, xpm, nu)]) + (-U[lk(g, xpn, mu)])) + (-U[lk(g, s, nu)]);
    return cos(ph); /* retrace (u1.hh:61-63) */
  }
}

/* retr_sum_Wplaquettes (flat-gauge_energy.hpp:31-94): sum_{x, mu<nu} eta ReTr P,
 * serial in the loop order (the reference's result at one OpenMP thread). */
ORC_API double orc_retr_sum_plaq(int grp, const int *dims, const double *U, double xi,
                                 int spatial_only, int compensated) {
  geom g = mkgeom(dims);
  const int aniso = (xi != 1.0);
  ksum acc = {0, 0};
  for (size_t s = 0; s < g.V; ++s) {
    long c[4];
    coords_of(&g, s, c);
    for (int mu = spatial_only ? 1 : 0; mu < g.nd - 1; ++mu)
      for (int nu = mu + 1; nu < g.nd; ++nu) {
        double v = plaq_retrace(grp, &g, U, c, s, mu, nu);
        if (aniso) {
          double eta = ((mu == 0) ^ (nu == 0)) ? 1.0 / xi : xi;
          v = eta * v;
        }
        kadd(&acc, v, compensated);
      }
  }
  return compensated ? kval(&acc) : acc.s;
}

/* get_S_G_hmc (flat-gaugemonomial.hh:46-54): -beta * sum / Nc */
ORC_API double orc_action(int grp, const int *dims, const double *U, double beta, double xi,
                          int compensated) {
  const double Nc = (grp == 1) ? 2.0 : 1.0;
  return -beta * orc_retr_sum_plaq(grp, dims, U, xi, 0, compensated) / Nc;
}

/* kineticmonomial (monomial.hh:33-47) with adjointfield dot (adjointfield.hh:219-238) */
ORC_API double orc_kinetic(int grp, const int *dims, const double *P, int compensated) {
  geom g = mkgeom(dims);
  size_t n = g.V * g.nd;
  ksum acc = {0, 0};
  for (size_t i = 0; i < n; ++i) {
    double v;
    if (grp == 1) v = P[3 * i] * P[3 * i] + P[3 * i + 1] * P[3 * i + 1] + P[3 * i + 2] * P[3 * i + 2];
    else v = P[i] * P[i];
    kadd(&acc, v, compensated);
  }
  return 0.5 * (compensated ? kval(&acc) : acc.s);
}

/* ------------------------------------------------------------------------ */
/* integrators (integrator.hh:15-241)                                       */
/* ------------------------------------------------------------------------ */

enum { ORC_LEAPFROG = 0, ORC_LP_LEAPFROG = 1, ORC_OMF4 = 2, ORC_LP_OMF4 = 3, ORC_EULER = 4,
       ORC_RUTH = 5, ORC_OMF2 = 6 };

typedef struct {
  int grp;
  geom g;
  double beta, xi;
  double *U, *P, *deriv;
  double rn; /* rounding multiplier for lp_* (0 = none) */
} mdctx;

static void P_(mdctx *m, double dt) { update_momenta(m->grp, &m->g, m->U, m->beta, m->xi, m->P, m->deriv, dt, m->rn); }
static void U_(mdctx *m, double dt) { update_gauge(m->grp, &m->g, m->U, m->P, dt, m->rn); }
static void restore_(mdctx *m) {
  int dims[5] = {(int)m->g.L[1], (int)m->g.L[2], (int)m->g.L[3], (int)m->g.L[0], m->g.nd};
  orc_restoreSU(m->grp, dims, m->U);
}

static void integrate(mdctx *m, int kind, long nsteps, double tau, int restore, long exponent) {
  const double dtau = tau / (double)nsteps;
  m->rn = 0.0;
  if (kind == ORC_LP_LEAPFROG || kind == ORC_LP_OMF4) m->rn = (double)(size_t)pow(10, (double)exponent);
  switch (kind) {
  case ORC_LEAPFROG:
  case ORC_LP_LEAPFROG: /* :26-50, :53-80 */
    P_(m, dtau / 2.);
    U_(m, dtau);
    for (long i = 0; i < nsteps - 1; ++i) { P_(m, dtau); U_(m, dtau); }
    P_(m, dtau / 2.);
    break;
  case ORC_OMF2: { /* :83-115 */
    const double lambda = 0.1938;
    const double om2l = (1. - 2. * lambda);
    P_(m, lambda * dtau);
    for (long i = 0; i < nsteps - 1; ++i) {
      U_(m, 0.5 * dtau);
      P_(m, om2l * dtau);
      U_(m, 0.5 * dtau);
      P_(m, 2 * lambda * dtau);
    }
    U_(m, 0.5 * dtau);
    P_(m, om2l * dtau);
    U_(m, 0.5 * dtau);
    P_(m, lambda * dtau);
    break;
  }
  case ORC_OMF4:
  case ORC_LP_OMF4: { /* :118-155, :158-198 (loop starts at i=1: n-1 steps) */
    const double rho = 0.2539785108410595, theta = -0.03230286765269967,
                 vartheta = 0.08398315262876693, lambda = 0.6822365335719091;
    const double eps[10] = {rho * dtau, lambda * dtau, theta * dtau,
                            0.5 * (1 - 2. * (lambda + vartheta)) * dtau,
                            (1 - 2. * (theta + rho)) * dtau,
                            0.5 * (1 - 2. * (lambda + vartheta)) * dtau,
                            theta * dtau, lambda * dtau, rho * dtau, 2 * vartheta * dtau};
    P_(m, 0.5 * eps[9]);
    for (long i = 1; i < nsteps - 1; ++i)
      for (int j = 0; j < 5; ++j) { U_(m, eps[2 * j]); P_(m, eps[2 * j + 1]); }
    for (int j = 0; j < 4; ++j) { U_(m, eps[2 * j]); P_(m, eps[2 * j + 1]); }
    U_(m, eps[8]);
    P_(m, 0.5 * eps[9]);
    break;
  }
  case ORC_EULER: /* :201-217 */
    for (long i = 0; i < nsteps; ++i) { P_(m, dtau); U_(m, dtau); }
    break;
  case ORC_RUTH: /* :220-241 */
    for (long i = 0; i < nsteps; ++i) {
      P_(m, dtau);
      U_(m, -1. / 24. * dtau);
      P_(m, -2. / 3. * dtau);
      U_(m, 3. / 4. * dtau);
      P_(m, 2. / 3. * dtau);
      U_(m, 7. / 24. * dtau);
    }
    break;
  default: /* set_integrator falls back to leapfrog (:274-277, :306-309) */
    integrate(m, ORC_LEAPFROG, nsteps, tau, 0, exponent);
    break;
  }
  if (restore) restore_(m);
}

ORC_API void orc_integrate(int grp, const int *dims, double *U, double *P, double beta, double xi,
                           int kind, long nsteps, double tau, int restore, long exponent) {
  mdctx m;
  m.grp = grp; m.g = mkgeom(dims); m.beta = beta; m.xi = xi; m.U = U; m.P = P;
  m.deriv = (double *)malloc(nmom(grp, &m.g) * sizeof(double));
  integrate(&m, kind, nsteps, tau, restore, exponent);
  free(m.deriv);
}

/* md_update (md_update.hh:18-81) with the monomial list {kinetic (md-passive),
 * gauge} of hmc.hpp:58-84.  out[] = {dH, accept, ddH, K_old, S_old, K_new, S_new,
 * uniform_drawn (or -1)}.  P_out (optional) receives the final momenta. */
ORC_API void orc_md_update(int grp, const int *dims, double *U, double beta, double xi, int kind,
                           long nsteps, double tau, long exponent, uint64_t seed, int revtest,
                           int compensated, double *out, double *P_out) {
  geom g = mkgeom(dims);
  size_t nm = nmom(grp, &g), nl = g.V * g.nd * (grp == 1 ? 4 : 1);
  orc_mt19937 e;
  orc_mt_seed(&e, seed);
  double *P = (double *)malloc(nm * sizeof(double));
  initnormal(grp, &g, P, &e);
  const double Kold = orc_kinetic(grp, dims, P, compensated);
  const double Sold = orc_action(grp, dims, U, beta, xi, compensated);
  double *Uold = (double *)malloc(nl * sizeof(double));
  memcpy(Uold, U, nl * sizeof(double));
  mdctx m;
  m.grp = grp; m.g = g; m.beta = beta; m.xi = xi; m.U = U; m.P = P;
  m.deriv = (double *)malloc(nm * sizeof(double));
  integrate(&m, kind, nsteps, tau, 1, exponent);
  const double Knew = orc_kinetic(grp, dims, P, compensated);
  const double Snew = orc_action(grp, dims, U, beta, xi, compensated);
  double dH = 0.;
  dH += Knew - Kold;
  dH += Snew - Sold;
  int accept = 1;
  double udraw = -1.0;
  if (dH > 0) {
    udraw = uniform(&e, 0.0, 1.0);
    if (udraw > exp(-dH)) accept = 0;
  }
  if (P_out) memcpy(P_out, P, nm * sizeof(double));
  double ddH = NAN;
  if (revtest) {
    double *Usave = (double *)malloc(nl * sizeof(double));
    memcpy(Usave, U, nl * sizeof(double));
    for (size_t i = 0; i < nm; ++i) P[i] = -P[i];
    integrate(&m, kind, nsteps, tau, 1, exponent);
    ddH = 0.;
    ddH += orc_kinetic(grp, dims, P, compensated) - Kold;
    ddH += orc_action(grp, dims, U, beta, xi, compensated) - Sold;
    memcpy(U, Usave, nl * sizeof(double));
    free(Usave);
  }
  if (!accept) memcpy(U, Uold, nl * sizeof(double));
  out[0] = dH; out[1] = accept; out[2] = ddH;
  out[3] = Kold; out[4] = Sold; out[5] = Knew; out[6] = Snew; out[7] = udraw;
  free(P); free(Uold); free(m.deriv);
}

/* ------------------------------------------------------------------------ */
/* clover energy density and topological charge (flat-energy_density.hh)    */
/* ------------------------------------------------------------------------ */

/* eps4 enumeration of tensors.hh:23-58 */
static void eps4(int idx[24][4], double val[24]) {
  int i = 0;
  for (int a = 1; a <= 4; a++)
    for (int b = 1; b <= 4; b++)
      for (int c = 1; c <= 4; c++)
        for (int d = 1; d <= 4; d++) {
          int p = (a - b) * (a - c) * (a - d) * (b - c) * (b - d) * (c - d);
          if (p != 0) {
            val[i] = p > 0 ? 1 : -1;
            idx[i][0] = a - 1; idx[i][1] = b - 1; idx[i][2] = c - 1; idx[i][3] = d - 1;
            i++;
          }
        }
}

/* energy_density (flat-energy_density.hh:57-164) */
ORC_API void orc_energy_density(int grp, const int *dims, const double *U, int clover, int ss,
                                double *E_out, double *Q_out) {
  geom g = mkgeom(dims);
  int eidx[24][4];
  double eval[24];
  eps4(eidx, eval);
  const int mu_start = ss ? 1 : 0;
  double res = 0., Q = 0.;
  for (size_t s = 0; s < g.V; ++s) {
    long c[4];
    coords_of(&g, s, c);
    su2 G[4][4];
    cplx Gc[4][4];
    memset(G, 0, sizeof(G));
    memset(Gc, 0, sizeof(Gc));
    for (int mu = mu_start; mu < g.nd - 1; ++mu)
      for (int nu = mu + 1; nu < g.nd; ++nu) {
        size_t x1 = nb(&g, c, mu, 1, -1, 0), x2 = nb(&g, c, nu, 1, -1, 0);
        if (grp == 1) {
#define L_(site, d) ld_su2(U, lk(&g, site, d))
          su2 leaf = su2_mul(su2_mul(su2_mul(L_(s, mu), L_(x1, nu)), su2_dag(L_(x2, mu))),
                             su2_dag(L_(s, nu)));
          if (clover) {
            size_t a1 = nb(&g, c, mu, -1, nu, 1), a2 = nb(&g, c, mu, -1, -1, 0);
            leaf = su2_add(leaf, su2_mul(su2_mul(su2_mul(L_(s, nu), su2_dag(L_(a1, mu))),
                                                 su2_dag(L_(a2, nu))), L_(a2, mu)));
            size_t b1 = nb(&g, c, mu, -1, -1, 0), b2 = nb(&g, c, mu, -1, nu, -1),
                   b3 = nb(&g, c, nu, -1, -1, 0);
            leaf = su2_add(leaf, su2_mul(su2_mul(su2_mul(su2_dag(L_(b1, mu)), su2_dag(L_(b2, nu))),
                                                 L_(b2, mu)), L_(b3, nu)));
            size_t c1 = nb(&g, c, nu, -1, -1, 0), c2 = nb(&g, c, nu, -1, mu, 1);
            leaf = su2_add(leaf, su2_mul(su2_mul(su2_mul(su2_dag(L_(c1, nu)), L_(c1, mu)),
                                                 L_(c2, nu)), su2_dag(L_(s, mu))));
          }
#undef L_
          /* traceless_antiherm (su2.hh:83-85): (0.5*(a-conj a), b) */
          su2 Gm = {0.5 * (leaf.ar - leaf.ar), 0.5 * (leaf.ai - (-leaf.ai)), leaf.br, leaf.bi};
          G[mu][nu] = Gm;
          res += 2. * su2_mul(Gm, Gm).ar;
        } else {
          double ph = ((U[lk(&g, s, mu)] + U[lk(&g, x1, nu)]) + (-U[lk(&g, x2, mu)])) + (-U[lk(&g, s, nu)]);
          cplx leaf = cexpi(ph);
          if (clover) {
            size_t a1 = nb(&g, c, mu, -1, nu, 1), a2 = nb(&g, c, mu, -1, -1, 0);
            leaf = cadd(leaf, cexpi(((U[lk(&g, s, nu)] + (-U[lk(&g, a1, mu)])) + (-U[lk(&g, a2, nu)])) + U[lk(&g, a2, mu)]));
            size_t b1 = nb(&g, c, mu, -1, -1, 0), b2 = nb(&g, c, mu, -1, nu, -1),
                   b3 = nb(&g, c, nu, -1, -1, 0);
            leaf = cadd(leaf, cexpi((((-U[lk(&g, b1, mu)]) + (-U[lk(&g, b2, nu)])) + U[lk(&g, b2, mu)]) + U[lk(&g, b3, nu)]));
            size_t c1 = nb(&g, c, nu, -1, -1, 0), c2 = nb(&g, c, nu, -1, mu, 1);
            leaf = cadd(leaf, cexpi((((-U[lk(&g, c1, nu)]) + U[lk(&g, c1, mu)]) + U[lk(&g, c2, nu)]) + (-U[lk(&g, s, mu)])));
          }
          /* traceless_antiherm(Complex) (u1.hh:81-83): (0, Im) */
          cplx Gm = {0., leaf.im};
          Gc[mu][nu] = Gm;
          res += cmul(Gm, Gm).re;
        }
      }
    if (g.nd == 4) {
      for (int i = 0; i < 24; ++i) {
        int i1 = eidx[i][0], i2 = eidx[i][1], i3 = eidx[i][2], i4 = eidx[i][3];
        if (i2 < i1) continue;
        if (i4 < i3) continue;
        double tr = (grp == 1) ? 2. * su2_mul(G[i1][i2], G[i3][i4]).ar
                               : cmul(Gc[i1][i2], Gc[i3][i4]).re;
        Q += eval[i] * tr;
      }
    }
    if (g.nd == 2) {
      if (grp == 1) {
        /* trace(G01 - G10) is real for SU(2) (su2.hh:77-80) */
        Q += -0.0;
      } else {
        cplx d = csub(Gc[0][1], Gc[1][0]);
        Q += -d.im;
      }
    }
  }
  res = -res / (double)g.V;
  if (clover) res /= 16.;
  Q = -4. * Q / (32.0 * M_PI * M_PI);
  if (clover) Q /= 16.;
  *E_out = res;
  *Q_out = Q;
}

/* ------------------------------------------------------------------------ */
/* Wilson gradient flow (flat-gradient_flow.hh:17-50, 70-197)               */
/* ------------------------------------------------------------------------ */

/* runge_kutta: Z = 0; for f: Z += 2 Nc z_f/beta * (beta/Nc) get_deriv(U K);
 * U <- exp(-eps e_f Z) U.  Z is caller scratch (momentum size). */
ORC_API void orc_runge_kutta(int grp, const int *dims, double *U, double *Z, double beta,
                             double xi, double eps) {
  geom g = mkgeom(dims);
  const double zfac[3] = {(-17.0) / (36.0), (8.0) / (9.0), (-3.0) / (4.0)};
  const double expfac[3] = {-36.0 / 4. / 17.0, 1., -1.};
  const double Nc = (grp == 1) ? 2.0 : 1.0;
  memset(Z, 0, nmom(grp, &g) * sizeof(double));
  for (int f = 0; f < 3; ++f) {
    orc_derivative(grp, dims, U, beta, xi, 2. * Nc * zfac[f] / beta, Z);
    update_gauge(grp, &g, U, Z, -eps * expfac[f], 0.0);
  }
}

/* gradient_flow driver (flat-gradient_flow.hh:70-197): rows of
 * {t, xi_R, P, P_ss, Ep, Ep_ss, Ec, Ec_ss, Q} written to rows[9*k]; returns the
 * number of rows (at most max_rows).  U is flowed in place. */
ORC_API long orc_gradient_flow(int grp, const int *dims, double *U, double beta, double xi,
                               double tmax, double eps, double tstart, double *rows,
                               long max_rows) {
  geom g = mkgeom(dims);
  const int d = g.nd;
  const double ndf = d * (d - 1) / 2.0, ndf_ss = (d - 1) * (d - 2) / 2.0;
  const double Nc = (grp == 1) ? 2.0 : 1.0;
  const double den = (double)g.V * Nc;
  double t[3], P[3] = {0}, E[3] = {0}, Q[3] = {0}, Pss[3] = {0}, Ess[3] = {0};
  for (int i = 0; i < 3; ++i) t[i] = tstart;
  double dens, tq, dens_ss, tq_ss;
  P[2] = orc_retr_sum_plaq(grp, dims, U, 1.0, 0, 0) / den;
  Pss[2] = orc_retr_sum_plaq(grp, dims, U, 1.0, 1, 0) / den;
  orc_energy_density(grp, dims, U, 1, 0, &dens, &tq);
  E[2] = dens;
  orc_energy_density(grp, dims, U, 1, 1, &dens_ss, &tq_ss);
  Ess[2] = dens_ss;
  double *Z = (double *)malloc(nmom(grp, &g) * sizeof(double));
  long nrow = 0;
  while (t[1] < tmax - eps) {
    t[0] = t[2]; P[0] = P[2]; E[0] = E[2]; Q[0] = Q[2]; Pss[0] = Pss[2]; Ess[0] = Ess[2];
    for (int k = 1; k < 3; ++k) {
      t[k] = t[k - 1] + eps;
      orc_runge_kutta(grp, dims, U, Z, beta, xi, eps);
      P[k] = orc_retr_sum_plaq(grp, dims, U, 1.0, 0, 0) / den;
      Pss[k] = orc_retr_sum_plaq(grp, dims, U, 1.0, 1, 0) / den;
      orc_energy_density(grp, dims, U, 1, 0, &dens, &tq);
      E[k] = dens;
      Q[k] = tq;
      orc_energy_density(grp, dims, U, 1, 1, &dens_ss, &tq_ss);
      Ess[k] = dens_ss;
    }
    const double Ep = ndf - P[1], Ep_ss = ndf_ss - Pss[1];
    const double Ep_ts = Ep - Ep_ss;
    const double xi2 = (Ep_ts / (d - 1)) / (Ep_ss / ndf_ss);
    if (nrow < max_rows) {
      double *r = rows + 9 * nrow;
      r[0] = t[1]; r[1] = sqrt(xi2); r[2] = P[1]; r[3] = Pss[1]; r[4] = Ep; r[5] = Ep_ss;
      r[6] = E[1]; r[7] = Ess[1]; r[8] = Q[1];
    }
    nrow++;
  }
  free(Z);
  return nrow < max_rows ? nrow : max_rows;
}

/* ------------------------------------------------------------------------ */
/* the HMC program loop (hmc.hpp:86-129,161-205; base_program.hpp:276-318)  */
/* ------------------------------------------------------------------------ */

/* Runs init (hotstart(seed, heat) + random_gauge_trafo(654321)) and n_meas
 * trajectories with engine mt19937(seed+i).  rows[7*i] = {i, accept, E*A, dH,
 * rho, ddH (NaN when not tested), Q}.  U holds the final configuration. */
ORC_API void orc_hmc_run(int grp, const int *dims, double *U, double beta, double xi,
                         uint64_t seed, int heat, int kind, long nsteps, double tau,
                         long exponent, long n_meas, long N_rev, double *rows) {
  geom g = mkgeom(dims);
  const double Nc = (grp == 1) ? 2.0 : 1.0;
  orc_hotstart(grp, dims, U, seed, heat ? 1.0 : 0.0);
  orc_random_gauge_trafo(grp, dims, U, 654321);
  const double fac = 2. / g.nd / (g.nd - 1);
  const double norm = fac / (double)g.V / Nc;
  double rate = 0.;
  for (long i = 0; i < n_meas; ++i) {
    int rev = (i > 0 && N_rev != 0 && (i % N_rev) == 0);
    double out[8];
    orc_md_update(grp, dims, U, beta, xi, kind, nsteps, tau, exponent, seed + (uint64_t)i, rev,
                  0, out, NULL);
    double energy = orc_retr_sum_plaq(grp, dims, U, 1.0, 0, 0);
    double E, Q;
    orc_energy_density(grp, dims, U, 1, 0, &E, &Q);
    rate += out[1];
    double *r = rows + 7 * i;
    r[0] = (double)i; r[1] = out[1]; r[2] = energy * norm; r[3] = out[0];
    r[4] = rate / (double)(i + 1); r[5] = rev ? out[2] : NAN; r[6] = Q;
  }
}

/* raw algebra entry points for the known-answer tests */
ORC_API void orc_su2_mul(const double *u, const double *v, double *out) {
  su2 a = {u[0], u[1], u[2], u[3]}, b = {v[0], v[1], v[2], v[3]};
  su2 r = su2_mul(a, b);
  out[0] = r.ar; out[1] = r.ai; out[2] = r.br; out[3] = r.bi;
}
ORC_API void orc_su2_restore(const double *u, double *out) {
  su2 a = {u[0], u[1], u[2], u[3]};
  su2 r = su2_restore(a);
  out[0] = r.ar; out[1] = r.ai; out[2] = r.br; out[3] = r.bi;
}
ORC_API void orc_su2_exp(const double *p, double *out) {
  su2 r = su2_exp(p[0], p[1], p[2]);
  out[0] = r.ar; out[1] = r.ai; out[2] = r.br; out[3] = r.bi;
}

/* ------------------------------------------------------------------------ */
/* Metropolis sweep                                                         */
/* ------------------------------------------------------------------------ */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
 * as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplications
 * (0xD2511F53, 0xCD9E8D57) with the Weyl key schedule (0x9E3779B9,
 * 0xBB67AE85).  Pinned by the paper's known-answer vectors (tests). */
static void philox4x32_10_c(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}
ORC_API void orc_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  philox4x32_10_c(c, key[0], key[1]);
  for (int k = 0; k < 4; ++k) out[k] = c[k];
}

static inline double u01_32(uint32_t w) { return ((double)w + 0.5) * (1.0 / 4294967296.0); }

/* the staple sum K of link (s, mu) (get_staples.hh:26-80): SU(2) in
 * derivative_link's order; U(1) with the anisotropy weight applied to the
 * exponential of the summed phases */
static su2 staples_su2(const geom *g, const double *U, size_t s, int mu, double xi, int aniso) {
  long c[4];
  coords_of(g, s, c);
  su2 K = {0, 0, 0, 0};
  for (int nu = 0; nu < g->nd; ++nu) {
    if (nu == mu) continue;
    double f = ((nu == 0) || (mu == 0)) ? 1.0 / xi : xi;
    su2 u1 = ld_su2(U, lk(g, nb(g, c, mu, 1, -1, 0), nu));
    if (aniso) u1 = su2_scale(f, u1);
    su2 t = su2_mul(su2_mul(u1, su2_dag(ld_su2(U, lk(g, nb(g, c, nu, 1, -1, 0), mu)))),
                    su2_dag(ld_su2(U, lk(g, s, nu))));
    K = su2_add(K, t);
  }
  for (int nu = 0; nu < g->nd; ++nu) {
    if (nu == mu) continue;
    double f = ((nu == 0) || (mu == 0)) ? 1.0 / xi : xi;
    su2 u1 = su2_dag(ld_su2(U, lk(g, nb(g, c, mu, 1, nu, -1), nu)));
    if (aniso) u1 = su2_scale(f, u1);
    size_t xm = nb(g, c, nu, -1, -1, 0);
    su2 t = su2_mul(su2_mul(u1, su2_dag(ld_su2(U, lk(g, xm, mu)))), ld_su2(U, lk(g, xm, nu)));
    K = su2_add(K, t);
  }
  return K;
}
static cplx staples_u1(const geom *g, const double *U, size_t s, int mu, double xi, int aniso) {
  long c[4];
  coords_of(g, s, c);
  cplx K = {0, 0};
  for (int nu = 0; nu < g->nd; ++nu) {
    if (nu == mu) continue;
    double th1 = U[lk(g, nb(g, c, mu, 1, -1, 0), nu)];
    double th2 = U[lk(g, nb(g, c, nu, 1, -1, 0), mu)];
    double th3 = U[lk(g, s, nu)];
    double f = (aniso && ((nu == 0) || (mu == 0))) ? 1.0 / xi : (aniso ? xi : 1.0);
    cplx t = cexpi((th1 + (-th2)) + (-th3));
    t.re *= f; t.im *= f;
    K = cadd(K, t);
  }
  for (int nu = 0; nu < g->nd; ++nu) {
    if (nu == mu) continue;
    size_t xm = nb(g, c, nu, -1, -1, 0);
    double th1 = U[lk(g, nb(g, c, mu, 1, nu, -1), nu)];
    double th2 = U[lk(g, xm, mu)];
    double th3 = U[lk(g, xm, nu)];
    double f = (aniso && ((nu == 0) || (mu == 0))) ? 1.0 / xi : (aniso ? xi : 1.0);
    cplx t = cexpi(((-th1) + (-th2)) + th3);
    t.re *= f; t.im *= f;
    K = cadd(K, t);
  }
  return K;
}

/* N_hit Metropolis hits on every link (flat-sweep.hh:64-164; one hit
 * :95-110): R = random_element(delta) (random_element.hh:9-34) from the
 * uniforms (w + 1/2) 2^-32 of one Philox4x32-10 block with counter
 * (s nd + mu, hit, sweep_lo, sweep_hi) and key (seed_lo, seed_hi) (alpha, u,
 * theta, acceptance; U(1): phase, -, -, acceptance), dS = beta/Nc (ReTr(U K) -
 * ReTr(U R K)), accept if dS < 0 or v < exp(-dS), U <- restoreSU(U R).  Links
 * are visited per direction mu, then colour class (sweep_colour: the site
 * parity where every extent is even), as the device sweep does; inside one
 * (mu, class) set the links are independent, so the site order does not
 * matter.  rates = {accepted / (N_hit
 * links), accepted temporal / (N_hit V)} (flat-sweep.hh:160-163). */
/* Independent sets of the sweep.  The staple of U_mu(x) holds, of the links
 * of direction mu, only U_mu(x +- nu) with nu != mu, so the links U_mu(x) of
 * one colour class of any proper colouring of the sites under the nu != mu
 * neighbour relation can be updated at once.  All extents in use even: the
 * two classes of the site parity (sum of the coordinates mod 2).  Otherwise
 * three classes, c(x) = sum_{nu != mu} h(x_nu) mod 3 with
 * h(k) = k + min(k, a) and a = (3 - L_nu mod 3) mod 3: each step along nu
 * changes h by 1 or 2 (mod 3), the wrap from L_nu - 1 to 0 by 1, so
 * neighbours never share a class. */
static int sweep_ncolours(const geom *g) {
  for (int d = 0; d < g->nd; ++d)
    if (g->L[d] % 2) return 3;
  return 2;
}
static int sweep_colour(const geom *g, const long c[4], int mu, int ncol) {
  if (ncol == 2) return (int)((c[0] + c[1] + c[2] + c[3]) & 1);
  long col = 0;
  for (int d = 0; d < g->nd; ++d) {
    if (d == mu) continue;
    const long a = (3 - g->L[d] % 3) % 3, k = c[d];
    col += k + (k < a ? k : a);
  }
  return (int)(col % 3);
}

ORC_API int orc_sweep_colour(const int *dims, const int *coords, int mu) {
  geom g = mkgeom(dims);
  const long c[4] = {coords[0], coords[1], coords[2], coords[3]};
  return sweep_colour(&g, c, mu, sweep_ncolours(&g));
}

ORC_API void orc_sweep(int grp, const int *dims, double *U, double beta, double xi, uint64_t seed,
                       uint64_t sweep, double delta, long n_hit, double *rates) {
  geom g = mkgeom(dims);
  const int aniso = (xi != 1.0);
  const double Nc = (grp == 1) ? 2.0 : 1.0;
  const double beta_nc = beta / Nc;
  const int ncol = sweep_ncolours(&g);
  double acc_all = 0, acc_t = 0;
  for (int mu = 0; mu < g.nd; ++mu) {
    for (int par = 0; par < ncol; ++par) {
      double acc_mu = 0;
#pragma omp parallel for schedule(static) reduction(+ : acc_mu)
      for (size_t s = 0; s < g.V; ++s) {
        long c[4];
        coords_of(&g, s, c);
        if (sweep_colour(&g, c, mu, ncol) != par) continue;
        const uint32_t lid = (uint32_t)lk(&g, s, mu);
        if (grp == 1) {
          const su2 K = staples_su2(&g, U, s, mu, xi, aniso);
          su2 u = ld_su2(U, lk(&g, s, mu));
          for (long n = 0; n < n_hit; ++n) {
            uint32_t w[4] = {lid, (uint32_t)n, (uint32_t)(sweep & 0xffffffffu), (uint32_t)(sweep >> 32)};
            philox4x32_10_c(w, (uint32_t)seed, (uint32_t)(seed >> 32));
            const double alpha = delta * (2.0 * ORC_PI) * u01_32(w[0]);
            const double uu = -1.0 + 2.0 * u01_32(w[1]);
            const double th = (2.0 * ORC_PI) * u01_32(w[2]);
            const double r = sqrt(1.0 - uu * uu);
            double sa, ca, st, ct;
            sincos(alpha, &sa, &ca);
            sincos(th, &st, &ct);
            const su2 R = {ca, sa * uu, (sa * r) * st, (sa * r) * ct};
            const su2 uR = su2_mul(u, R);
            const double dS = beta_nc * (2.0 * su2_mul(u, K).ar - 2.0 * su2_mul(uR, K).ar);
            if (dS < 0.0 || u01_32(w[3]) < exp(-dS)) {
              u = su2_restore(uR);
              acc_mu += 1.0;
            }
          }
          st_su2(U, lk(&g, s, mu), u);
        } else {
          const cplx K = staples_u1(&g, U, s, mu, xi, aniso);
          double u = U[lk(&g, s, mu)];
          for (long n = 0; n < n_hit; ++n) {
            uint32_t w[4] = {lid, (uint32_t)n, (uint32_t)(sweep & 0xffffffffu), (uint32_t)(sweep >> 32)};
            philox4x32_10_c(w, (uint32_t)seed, (uint32_t)(seed >> 32));
            const double th = -ORC_PI * delta + (2.0 * ORC_PI * delta) * u01_32(w[0]);
            const double un = u + th;
            const double dS = beta_nc * (cmul(cexpi(u), K).re - cmul(cexpi(un), K).re);
            if (dS < 0.0 || u01_32(w[3]) < exp(-dS)) {
              u = un;
              acc_mu += 1.0;
            }
          }
          U[lk(&g, s, mu)] = u;
        }
      }
      acc_all += acc_mu;
      if (mu == 0) acc_t += acc_mu;
    }
  }
  const double nh = n_hit > 0 ? (double)n_hit : 1.0;
  rates[0] = acc_all / nh / (double)(g.V * g.nd);
  rates[1] = acc_t / nh / (double)g.V;
}
