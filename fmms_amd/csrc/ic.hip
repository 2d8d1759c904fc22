// ic.hip — seeded initial conditions (host code; G = 1, total mass 1).
//
// Extends the samplers of models.rs:35-82 with velocities and the clustered
// and disk models of the benchmark workloads:
//   * truncated Plummer with isotropic distribution-function velocities
//     (Aarseth, Henon & Wielen 1974 rejection on q^2 (1 - q^2)^{7/2});
//   * truncated Hernquist with a local Maxwellian of the isotropic Jeans
//     dispersion (Hernquist 1990 eq. 10), speeds kept below 0.95 v_esc;
//   * exponential disk (Gamma(2) radii, sech^2 heights) on circular orbits
//     with v_c^2 = m_c / R + Freeman's thin-disk term;
//   * Hernquist halos of log-uniform masses and scale radii in a uniform
//     unit-cube background.
// Positions of the Plummer and Hernquist models follow models.rs:35-82 draw
// for draw (the Plummer positions equal fmmh_plummer_trunc's); velocities come
// from a second SplitMix64 stream so they never perturb the positions.
// Compiled with -ffp-contract=off: every value is reproducible bit for bit.
#include <cmath>
#include <cstdint>
#include <vector>

#include "fmmh.h"

namespace {

struct Rng {  // models.rs:13-31 SplitMix64
  uint64_t s;
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double unit() { return (double)(next() >> 11) / (double)(1ull << 53); }
  // Box-Muller as energy_gate.rs:31-35 (u1 floored at 1e-15)
  double gauss() {
    const double u1 = std::fmax(unit(), 1e-15);
    const double u2 = unit();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * PI * u2);
  }
  static constexpr double PI = 3.14159265358979323846;
};

// Independent stream for a (seed, tag) pair: the SplitMix64 output mix of
// seed + tag * (odd constant), so streams start at unrelated states.
uint64_t stream_seed(uint64_t seed, uint64_t tag) {
  uint64_t z = seed + (tag + 1) * 0xD1B54A32D192ED03ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

constexpr uint64_t TAG_VEL = 1, TAG_HALO_PARAMS = 2, TAG_BACKGROUND = 3, TAG_HALO0 = 16;

void direction(Rng& r, double len, double* out) {  // models.rs:46-55
  const double z = 2.0 * r.unit() - 1.0;
  const double phi = 2.0 * Rng::PI * r.unit();
  const double s = std::sqrt(1.0 - z * z);
  double sp, cp;
  ::sincos(phi, &sp, &cp);
  out[0] = len * s * cp;
  out[1] = len * s * sp;
  out[2] = len * z;
}

// Plummer radius (models.rs:41-44), redrawn until r <= rmax.
double plummer_radius(Rng& r, double rmax) {
  double rad;
  do {
    const double u = r.unit();
    const double w = std::pow(u, 2.0 / 3.0);
    rad = std::sqrt(w / (1.0 - w));
  } while (!(rad <= rmax));
  return rad;
}

// Hernquist radius for scale a (models.rs:68-69), redrawn until r <= rmax.
double hernquist_radius(Rng& r, double a, double rmax) {
  double rad;
  do {
    const double s = std::sqrt(r.unit());
    rad = a * (s / (1.0 - s));
  } while (!(rad <= rmax));
  return rad;
}

// Isotropic 1-D dispersion of a Hernquist sphere (G = M = 1, scale a) at r:
// Hernquist 1990 eq. 10, and for r >= 6a its expansion in a/r (20 terms,
// exact to double precision there; the closed form cancels badly at large r).
double hernquist_sigma2(double r, double a) {
  const double u = r / a;
  if (u >= 6.0) {
    static const double c[20] = {
        12.0 / 5,       -14.0 / 5,      102.0 / 35,     -207.0 / 70,    125.0 / 42,
        -209.0 / 70,    329.0 / 110,    -494.0 / 165,   2142.0 / 715,   -3000.0 / 1001,
        1364.0 / 455,   -5457.0 / 1820, 7137.0 / 2380,  -3059.0 / 1020, 3875.0 / 1292,
        -4844.0 / 1615, 5984.0 / 1995,  -21942.0 / 7315, 26562.0 / 8855, -10625.0 / 3542};
    const double x = 1.0 / u;
    double s = 0.0;
    for (int k = 19; k >= 0; k--) s = s * x + c[k];
    return s * x / (12.0 * a);
  }
  const double u1 = 1.0 + u;
  const double f = 12.0 * u * u1 * u1 * u1 * std::log(u1 / u) -
                   u / u1 * (25.0 + 52.0 * u + 42.0 * u * u + 12.0 * u * u * u);
  return f / (12.0 * a);
}

void remove_mean_velocity(size_t n, const double* mass, double* vel) {
  double p[3] = {0.0, 0.0, 0.0}, mt = 0.0;
  for (size_t i = 0; i < n; i++) {
    mt += mass[i];
    for (int d = 0; d < 3; d++) p[d] += mass[i] * vel[3 * i + d];
  }
  if (!(mt > 0.0)) return;
  for (int d = 0; d < 3; d++) p[d] /= mt;
  for (size_t i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) vel[3 * i + d] -= p[d];
}

// Maxwellian velocity of dispersion^2 s2, redrawn while |v| >= vmax.
void maxwell(Rng& r, double s2, double vmax, double* v) {
  const double s = std::sqrt(s2);
  for (;;) {
    v[0] = s * r.gauss();
    v[1] = s * r.gauss();
    v[2] = s * r.gauss();
    if (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] < vmax * vmax) return;
  }
}

// One Hernquist system of mass mh, scale a, truncated at rmax, about c.
void hernquist_bodies(Rng& rp, Rng& rv, size_t n, double mh, double a, double rmax, const double* c,
                      double mb, double* pos, double* vel, double* mass) {
  for (size_t i = 0; i < n; i++) {
    const double r = hernquist_radius(rp, a, rmax);
    double x[3];
    direction(rp, r, x);
    for (int d = 0; d < 3; d++) pos[3 * i + d] = c[d] + x[d];
    const double s2 = mh * hernquist_sigma2(r, a);
    const double vesc = std::sqrt(2.0 * mh / (r + a));
    maxwell(rv, s2, 0.95 * vesc, vel + 3 * i);
    mass[i] = mb;
  }
}

}  // namespace

extern "C" {

int fmmh_ic_plummer(size_t n, double rmax, uint64_t seed, double* pos, double* vel, double* mass) {
  if (n == 0 || !pos || !vel || !mass || !(rmax > 0.0)) return FMMH_E_PARAM;
  Rng rp{seed}, rv{stream_seed(seed, TAG_VEL)};
  const double m = 1.0 / (double)n;
  for (size_t i = 0; i < n; i++) {
    const double r = plummer_radius(rp, rmax);
    direction(rp, r, pos + 3 * i);
    mass[i] = m;
    // q = v / v_esc from g(q) = q^2 (1 - q^2)^{7/2} (max 0.0920 < 0.1)
    double q;
    for (;;) {
      const double x1 = rv.unit(), x2 = rv.unit();
      if (0.1 * x2 < x1 * x1 * std::pow(1.0 - x1 * x1, 3.5)) {
        q = x1;
        break;
      }
    }
    const double vesc = std::sqrt(2.0) * std::pow(1.0 + r * r, -0.25);
    direction(rv, q * vesc, vel + 3 * i);
  }
  remove_mean_velocity(n, mass, vel);
  return FMMH_OK;
}

int fmmh_ic_hernquist(size_t n, double a, double rmax, uint64_t seed, double* pos, double* vel,
                      double* mass) {
  if (n == 0 || !pos || !vel || !mass || !(a > 0.0) || !(rmax > 0.0)) return FMMH_E_PARAM;
  Rng rp{seed}, rv{stream_seed(seed, TAG_VEL)};
  const double c[3] = {0.0, 0.0, 0.0};
  hernquist_bodies(rp, rv, n, 1.0, a, rmax, c, 1.0 / (double)n, pos, vel, mass);
  remove_mean_velocity(n, mass, vel);
  return FMMH_OK;
}

int fmmh_ic_disk(size_t n, double rd, double zd, double rmax, double m_central, double sigma,
                 uint64_t seed, double* pos, double* vel, double* mass) {
  if (n == 0 || !pos || !vel || !mass || !(rd > 0.0) || !(zd > 0.0) || !(rmax > 0.0) ||
      !(m_central >= 0.0 && m_central < 1.0) || !(sigma >= 0.0) || (m_central > 0.0 && n < 2))
    return FMMH_E_PARAM;
  Rng rp{seed}, rv{stream_seed(seed, TAG_VEL)};
  size_t i0 = 0;
  if (m_central > 0.0) {  // central point mass at rest at the origin
    for (int d = 0; d < 3; d++) pos[d] = vel[d] = 0.0;
    mass[0] = m_central;
    i0 = 1;
  }
  const double md = 1.0 - m_central;
  const double mb = md / (double)(n - i0);
  for (size_t i = i0; i < n; i++) {
    double R;
    do {  // surface density exp(-R/rd): R ~ Gamma(2, rd)
      R = -rd * std::log((1.0 - rp.unit()) * (1.0 - rp.unit()));
    } while (!(R <= rmax));
    const double phi = 2.0 * Rng::PI * rp.unit();
    double w;
    do {
      w = 2.0 * rp.unit() - 1.0;
    } while (!(w > -1.0));
    const double z = zd * std::atanh(w);  // density sech^2(z / zd)
    double sp, cp;
    ::sincos(phi, &sp, &cp);
    pos[3 * i] = R * cp;
    pos[3 * i + 1] = R * sp;
    pos[3 * i + 2] = z;
    mass[i] = mb;
    double vc2 = 0.0;
    if (R > 0.0) {
      const double y = R / (2.0 * rd);
      const double fr = std::cyl_bessel_i(0.0, y) * std::cyl_bessel_k(0.0, y) -
                        std::cyl_bessel_i(1.0, y) * std::cyl_bessel_k(1.0, y);
      vc2 = m_central / R + 2.0 * md / rd * y * y * fr;
    }
    const double vc = std::sqrt(vc2);
    vel[3 * i] = -vc * sp;
    vel[3 * i + 1] = vc * cp;
    vel[3 * i + 2] = 0.0;
    if (sigma > 0.0)
      for (int d = 0; d < 3; d++) vel[3 * i + d] += sigma * rv.gauss();
  }
  remove_mean_velocity(n, mass, vel);
  return FMMH_OK;
}

int fmmh_ic_halos_cube(size_t n, int nhalo, double halo_frac, double a_min, double a_max,
                       double mass_decades, double rmax_over_a, uint64_t seed, double* pos,
                       double* vel, double* mass) {
  if (n == 0 || !pos || !vel || !mass || nhalo < 0 || !(halo_frac >= 0.0 && halo_frac <= 1.0) ||
      !(a_min > 0.0) || !(a_max >= a_min) || !(mass_decades >= 0.0) || !(rmax_over_a > 0.0) ||
      (nhalo == 0 && halo_frac > 0.0))
    return FMMH_E_PARAM;
  // halo parameters: log-uniform masses and scale radii, centres in [0,1)^3
  Rng rh{stream_seed(seed, TAG_HALO_PARAMS)};
  std::vector<double> w(nhalo), a(nhalo), c(3 * (size_t)nhalo);
  double wsum = 0.0;
  for (int h = 0; h < nhalo; h++) {
    w[h] = std::pow(10.0, mass_decades * rh.unit());
    a[h] = a_min * std::pow(a_max / a_min, rh.unit());
    for (int d = 0; d < 3; d++) c[3 * h + d] = rh.unit();
    wsum += w[h];
  }
  // body counts by largest remainder of halo_frac * n * w / wsum
  std::vector<size_t> cnt(nhalo);
  const size_t nh_tot = (size_t)std::floor(halo_frac * (double)n + 0.5);
  size_t given = 0;
  std::vector<double> rem(nhalo);
  for (int h = 0; h < nhalo; h++) {
    const double x = (double)nh_tot * w[h] / wsum;
    cnt[h] = (size_t)std::floor(x);
    rem[h] = x - (double)cnt[h];
    given += cnt[h];
  }
  while (given < nh_tot) {  // one extra body to the largest remainders, first index on ties
    int best = 0;
    for (int h = 1; h < nhalo; h++)
      if (rem[h] > rem[best]) best = h;
    cnt[best]++;
    rem[best] = -1.0;
    given++;
  }
  const double mb = 1.0 / (double)n;
  size_t off = 0;
  for (int h = 0; h < nhalo; h++) {
    Rng rp{stream_seed(seed, TAG_HALO0 + 2 * (uint64_t)h)};
    Rng rv{stream_seed(seed, TAG_HALO0 + 2 * (uint64_t)h + 1)};
    const double mh = (double)cnt[h] * mb;
    hernquist_bodies(rp, rv, cnt[h], mh, a[h], rmax_over_a * a[h], &c[3 * h], mb, pos + 3 * off,
                     vel + 3 * off, mass + off);
    off += cnt[h];
  }
  Rng rb{stream_seed(seed, TAG_BACKGROUND)};
  for (size_t i = off; i < n; i++) {  // cold uniform background
    for (int d = 0; d < 3; d++) pos[3 * i + d] = rb.unit();
    for (int d = 0; d < 3; d++) vel[3 * i + d] = 0.0;
    mass[i] = mb;
  }
  return FMMH_OK;
}

}  // extern "C"
