// ic_oracle.cpp — CPU restatement of the initial-condition library and of
// the energy bookkeeping of crates/fmm/tests/energy_gate.rs.
//
// TEST INFRASTRUCTURE ONLY (see fmm_oracle.h). Compiled with
// -ffp-contract=off like the rest of the oracle.
//
// The generators are extensions of models.rs:35-82; their distributions are
// pinned in tests/test_ic.py against closed forms and numerical integrals
// (Plummer speed distribution and virial ratio, the Hernquist Jeans integral,
// Freeman's disk curve against its Hankel-transform integral), and the
// product must reproduce these draws bit for bit.
#include <cmath>
#include <cstdint>
#include <cstddef>
#include <vector>

#include "fmm_oracle.h"

namespace {

const double PI = 3.14159265358979323846;

uint64_t mix64(uint64_t z) {  // SplitMix64 output function (models.rs:24-27)
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Rng {  // models.rs:13-31
  uint64_t s;
  double unit() {
    s += 0x9E3779B97F4A7C15ull;
    return (double)(mix64(s) >> 11) / (double)(1ull << 53);
  }
};

double gaussian(Rng& r) {  // energy_gate.rs:31-35
  double u1 = r.unit();
  if (u1 < 1e-15) u1 = 1e-15;
  double u2 = r.unit();
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * PI * u2);
}

// seeds of the independent streams: mix64(seed + (tag + 1) * C)
uint64_t sub_seed(uint64_t seed, uint64_t tag) { return mix64(seed + (tag + 1) * 0xD1B54A32D192ED03ull); }

void unit_vector_scaled(Rng& r, double len, double* v) {  // models.rs:46-55
  double z = 2.0 * r.unit() - 1.0;
  double phi = 2.0 * PI * r.unit();
  double s = std::sqrt(1.0 - z * z);
  double sn, cs;
  sincos(phi, &sn, &cs);
  v[0] = len * s * cs;
  v[1] = len * s * sn;
  v[2] = len * z;
}

double sigma2_hernquist(double r, double a) {
  double u = r / a;
  if (u >= 6.0) {  // series of eq. 10 in x = a/r
    const double num[20] = {12, -14, 102, -207, 125, -209, 329, -494, 2142, -3000,
                            1364, -5457, 7137, -3059, 3875, -4844, 5984, -21942, 26562, -10625};
    const double den[20] = {5, 5, 35, 70, 42, 70, 110, 165, 715, 1001,
                            455, 1820, 2380, 1020, 1292, 1615, 1995, 7315, 8855, 3542};
    double x = 1.0 / u, acc = 0.0;
    for (int k = 19; k >= 0; k--) acc = acc * x + num[k] / den[k];
    return acc * x / (12.0 * a);
  }
  double v = 1.0 + u;  // Hernquist 1990 eq. 10 with G = M = 1
  double t1 = 12.0 * u * v * v * v * std::log(v / u);
  double t2 = u / v * (25.0 + 52.0 * u + 42.0 * u * u + 12.0 * u * u * u);
  return (t1 - t2) / (12.0 * a);
}

void zero_momentum(size_t n, const double* m, double* v) {
  double px = 0, py = 0, pz = 0, mt = 0;
  for (size_t i = 0; i < n; i++) {
    mt += m[i];
    px += m[i] * v[3 * i];
    py += m[i] * v[3 * i + 1];
    pz += m[i] * v[3 * i + 2];
  }
  if (!(mt > 0.0)) return;
  px /= mt;
  py /= mt;
  pz /= mt;
  for (size_t i = 0; i < n; i++) {
    v[3 * i] -= px;
    v[3 * i + 1] -= py;
    v[3 * i + 2] -= pz;
  }
}

void hernquist_system(Rng& rp, Rng& rv, size_t n, double mh, double a, double rmax,
                      const double* c, double mb, double* pos, double* vel, double* mass) {
  for (size_t i = 0; i < n; i++) {
    double r;
    for (;;) {
      double s = std::sqrt(rp.unit());
      r = a * (s / (1.0 - s));
      if (r <= rmax) break;
    }
    double d[3];
    unit_vector_scaled(rp, r, d);
    pos[3 * i] = c[0] + d[0];
    pos[3 * i + 1] = c[1] + d[1];
    pos[3 * i + 2] = c[2] + d[2];
    double sg = std::sqrt(mh * sigma2_hernquist(r, a));
    double vmax = 0.95 * std::sqrt(2.0 * mh / (r + a));
    for (;;) {
      double vx = sg * gaussian(rv), vy = sg * gaussian(rv), vz = sg * gaussian(rv);
      vel[3 * i] = vx;
      vel[3 * i + 1] = vy;
      vel[3 * i + 2] = vz;
      if (vx * vx + vy * vy + vz * vz < vmax * vmax) break;
    }
    mass[i] = mb;
  }
}

}  // namespace

extern "C" {

double orc_hernquist_sigma2(double r, double a) { return sigma2_hernquist(r, a); }

int orc_ic_plummer(size_t n, double rmax, uint64_t seed, double* pos, double* vel, double* mass) {
  if (n == 0 || !(rmax > 0.0)) return 2;
  Rng rp{seed}, rv{sub_seed(seed, 1)};
  for (size_t i = 0; i < n; i++) {
    double r;
    for (;;) {  // models.rs:41-44, redrawn beyond rmax
      double w = std::pow(rp.unit(), 2.0 / 3.0);
      r = std::sqrt(w / (1.0 - w));
      if (r <= rmax) break;
    }
    unit_vector_scaled(rp, r, pos + 3 * i);
    mass[i] = 1.0 / (double)n;
    double q;  // Aarseth, Henon & Wielen (1974): von Neumann rejection under 0.1
    do {
      q = rv.unit();
    } while (!(0.1 * rv.unit() < q * q * std::pow(1.0 - q * q, 3.5)));
    double ve = std::sqrt(2.0) * std::pow(1.0 + r * r, -0.25);
    unit_vector_scaled(rv, q * ve, vel + 3 * i);
  }
  zero_momentum(n, mass, vel);
  return 0;
}

int orc_ic_hernquist(size_t n, double a, double rmax, uint64_t seed, double* pos, double* vel,
                     double* mass) {
  if (n == 0 || !(a > 0.0) || !(rmax > 0.0)) return 2;
  Rng rp{seed}, rv{sub_seed(seed, 1)};
  const double c[3] = {0, 0, 0};
  hernquist_system(rp, rv, n, 1.0, a, rmax, c, 1.0 / (double)n, pos, vel, mass);
  zero_momentum(n, mass, vel);
  return 0;
}

int orc_ic_disk(size_t n, double rd, double zd, double rmax, double mc, double sigma,
                uint64_t seed, double* pos, double* vel, double* mass) {
  if (n == 0 || !(rd > 0) || !(zd > 0) || !(rmax > 0) || !(mc >= 0 && mc < 1) || !(sigma >= 0) ||
      (mc > 0 && n < 2))
    return 2;
  Rng rp{seed}, rv{sub_seed(seed, 1)};
  size_t first = mc > 0.0 ? 1 : 0;
  if (first) {
    pos[0] = pos[1] = pos[2] = 0.0;
    vel[0] = vel[1] = vel[2] = 0.0;
    mass[0] = mc;
  }
  double md = 1.0 - mc;
  for (size_t i = first; i < n; i++) {
    double R;
    for (;;) {
      double u1 = 1.0 - rp.unit();
      double u2 = 1.0 - rp.unit();
      R = -rd * std::log(u1 * u2);
      if (R <= rmax) break;
    }
    double phi = 2.0 * PI * rp.unit();
    double w;
    do {
      w = 2.0 * rp.unit() - 1.0;
    } while (w <= -1.0);
    double sn, cs;
    sincos(phi, &sn, &cs);
    pos[3 * i] = R * cs;
    pos[3 * i + 1] = R * sn;
    pos[3 * i + 2] = zd * std::atanh(w);
    mass[i] = md / (double)(n - first);
    double vc = 0.0;
    if (R > 0.0) {  // point mass + Freeman (1970) thin exponential disk
      double y = R / (2.0 * rd);
      double b = std::cyl_bessel_i(0.0, y) * std::cyl_bessel_k(0.0, y) -
                 std::cyl_bessel_i(1.0, y) * std::cyl_bessel_k(1.0, y);
      vc = std::sqrt(mc / R + 2.0 * md / rd * y * y * b);
    }
    vel[3 * i] = -vc * sn;
    vel[3 * i + 1] = vc * cs;
    vel[3 * i + 2] = 0.0;
    if (sigma > 0.0) {
      vel[3 * i] += sigma * gaussian(rv);
      vel[3 * i + 1] += sigma * gaussian(rv);
      vel[3 * i + 2] += sigma * gaussian(rv);
    }
  }
  zero_momentum(n, mass, vel);
  return 0;
}

int orc_ic_halos_cube(size_t n, int nh, double frac, double amin, double amax, double decades,
                      double rmax_a, uint64_t seed, double* pos, double* vel, double* mass) {
  if (n == 0 || nh < 0 || !(frac >= 0 && frac <= 1) || !(amin > 0) || !(amax >= amin) ||
      !(decades >= 0) || !(rmax_a > 0) || (nh == 0 && frac > 0))
    return 2;
  Rng rh{sub_seed(seed, 2)};
  std::vector<double> wt(nh), sc(nh), cen(3 * (size_t)nh);
  double wtot = 0.0;
  for (int h = 0; h < nh; h++) {
    wt[h] = std::pow(10.0, decades * rh.unit());
    sc[h] = amin * std::pow(amax / amin, rh.unit());
    cen[3 * h] = rh.unit();
    cen[3 * h + 1] = rh.unit();
    cen[3 * h + 2] = rh.unit();
    wtot += wt[h];
  }
  size_t total = (size_t)std::floor(frac * (double)n + 0.5);
  std::vector<size_t> cnt(nh);
  std::vector<double> frac_part(nh);
  size_t assigned = 0;
  for (int h = 0; h < nh; h++) {
    double share = (double)total * wt[h] / wtot;
    cnt[h] = (size_t)std::floor(share);
    frac_part[h] = share - (double)cnt[h];
    assigned += cnt[h];
  }
  for (; assigned < total; assigned++) {  // largest remainders, lowest index on ties
    int k = 0;
    for (int h = 1; h < nh; h++)
      if (frac_part[h] > frac_part[k]) k = h;
    cnt[k] += 1;
    frac_part[k] = -1.0;
  }
  size_t at = 0;
  for (int h = 0; h < nh; h++) {
    Rng rp{sub_seed(seed, 16 + 2 * (uint64_t)h)}, rv{sub_seed(seed, 17 + 2 * (uint64_t)h)};
    hernquist_system(rp, rv, cnt[h], (double)cnt[h] * (1.0 / (double)n), sc[h], rmax_a * sc[h],
                     &cen[3 * h], 1.0 / (double)n, pos + 3 * at, vel + 3 * at, mass + at);
    at += cnt[h];
  }
  Rng rb{sub_seed(seed, 3)};
  for (size_t i = at; i < n; i++) {
    pos[3 * i] = rb.unit();
    pos[3 * i + 1] = rb.unit();
    pos[3 * i + 2] = rb.unit();
    vel[3 * i] = vel[3 * i + 1] = vel[3 * i + 2] = 0.0;
    mass[i] = 1.0 / (double)n;
  }
  return 0;
}

// energy_gate.rs:39-79 warm start: sigma^2(r) = 1/(6 sqrt(1 + r^2)) Maxwellian
// from Rng(43), centre-of-mass drift removed, rescaled to 2T = -W with
// W = -1/2 sum m pot.
void orc_warm_start(size_t n, const double* pos, const double* mass, const double* pot,
                    double* vel) {
  Rng rng{43};
  for (size_t i = 0; i < n; i++) {
    double r2 = pos[3 * i] * pos[3 * i] + pos[3 * i + 1] * pos[3 * i + 1] +
                pos[3 * i + 2] * pos[3 * i + 2];
    double sigma = std::sqrt(1.0 / (6.0 * std::sqrt(1.0 + r2)));
    for (int d = 0; d < 3; d++) vel[3 * i + d] = sigma * gaussian(rng);
  }
  double mtot = 0.0;
  for (size_t i = 0; i < n; i++) mtot += mass[i];
  double pc[3] = {0, 0, 0};
  for (size_t i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) pc[d] += mass[i] * vel[3 * i + d];
  for (size_t i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) vel[3 * i + d] -= pc[d] / mtot;
  double t0 = orc_kinetic(n, mass, vel);
  double w0 = orc_potential_energy(n, mass, pot);
  if (t0 > 0.0) {
    double c = std::sqrt(-0.5 * w0 / t0);
    for (size_t i = 0; i < 3 * n; i++) vel[i] *= c;
  }
}

double orc_kinetic(size_t n, const double* mass, const double* vel) {  // energy_gate.rs:90-97
  double t = 0.0;
  for (size_t i = 0; i < n; i++) {
    const double* v = vel + 3 * i;
    t += 0.5 * mass[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  }
  return t;
}

double orc_potential_energy(size_t n, const double* mass, const double* pot) {  // :100-106
  double w = 0.0;
  for (size_t i = 0; i < n; i++) w += mass[i] * pot[i];
  return -0.5 * w;
}

double orc_momentum_rel(size_t n, const double* mass, const double* vel) {  // :109-123
  double p[3] = {0, 0, 0}, mag = 0.0;
  for (size_t i = 0; i < n; i++) {
    const double* v = vel + 3 * i;
    double vm = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int d = 0; d < 3; d++) p[d] += mass[i] * v[d];
    mag += mass[i] * vm;
  }
  double norm = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  return norm / (mag > 2.2250738585072014e-308 ? mag : 2.2250738585072014e-308);
}

}  // extern "C"
