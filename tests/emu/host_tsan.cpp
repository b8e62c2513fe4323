// TEST INFRASTRUCTURE ONLY: the commit's host code (sif_amd/csrc/
// host_commit.h, rung_voids.h, batch_pairs.h) driven the way the exodus
// driver drives it, built for the CPU with a sanitizer (thread, or address
// and undefined behaviour) and run as a program.
//
// Per round, as the driver's batch loop does:
//   - the commit, on this thread, accepts some voids into the catalogue
//     (left unlinked: the cell list is linked when next read);
//   - the next batch's assembly links the list, then tests a block of
//     candidates on the pool with the read-only scan (edge candidates, whose
//     scan wraps the box, included);
//   - the batch's pair effects are listed on the pool.
// Every pool answer is checked against the same call made serially; a read of
// an unlinked list must be counted. Exit status 0 and "host harness ok" on
// success.
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <vector>

#include "batch_pairs.h"
#include "host_commit.h"

using namespace sif_amd;

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 1) {}
  uint64_t next() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
  }
  float uniform(float lo, float hi) { return lo + (hi - lo) * (float)((next() >> 40) * 0x1p-24); }
};

int failures = 0;
#define EXPECT(cond, ...)                   \
  do {                                      \
    if (!(cond)) {                          \
      fprintf(stderr, "FAIL: " __VA_ARGS__); \
      fprintf(stderr, "\n");                \
      failures++;                           \
    }                                       \
  } while (0)

}  // namespace

int main(int argc, char** argv) {
  const unsigned workers = argc > 1 ? (unsigned)atoi(argv[1]) : 7;
  const int rounds = argc > 2 ? atoi(argv[2]) : 40;
  const float box = 250.0f, f = 0.3f, scale = 1.0f;
  const uint32_t n = 64;
  const float l = box / (float)n;
  WorkPool pool(workers);
  Rng rng(7);

  Voids V;
  V.init(box);
  RungVoids BI;
  BI.init(box);
  BI.columns = false;
  std::vector<uint8_t> got;
  std::vector<uint32_t> place;

  for (int round = 0; round < rounds; round++) {
    // the commit: a few voids, some straddling the faces
    const int add = 5 + (int)(rng.next() % 20);
    for (int a = 0; a < add; a++) {
      float x = rng.uniform(0, box), y = rng.uniform(0, box), z = rng.uniform(0, box);
      if (a % 3 == 0)
        x = rng.uniform(0, 1) < 0.5f ? rng.uniform(0, 4) : box - rng.uniform(0, 4);
      V.add(x, y, z, rng.uniform(1.0f, 18.0f), 0);
    }
    // the next batch's assembly: link once, then the read-only scan on the pool
    const uint64_t m = 3000;
    std::vector<float> q(4 * m);
    for (uint64_t t = 0; t < m; t++) {
      // candidates on nodes, a third of them within a search radius of a face
      uint32_t ix = (uint32_t)(rng.next() % n), iy = (uint32_t)(rng.next() % n),
               iz = (uint32_t)(rng.next() % n);
      if (t % 3 == 0)
        iz = (uint32_t)(rng.next() % 3);
      q[4 * t] = (float)ix * l;
      q[4 * t + 1] = (float)iy * l;
      q[4 * t + 2] = (float)iz * l;
      q[4 * t + 3] = rng.uniform(2.0f, 12.0f);
    }
    got.assign(m, 0);
    V.link();
    pool.run(m, [&](uint64_t a, uint64_t b) {
      for (uint64_t t = a; t < b; t++)
        got[t] = V.overlap_linked(q[4 * t], q[4 * t + 1], q[4 * t + 2], q[4 * t + 3], f, scale);
    });
    for (uint64_t t = 0; t < m; t++)
      EXPECT(got[t] == (uint8_t)V.overlap_exact(q[4 * t], q[4 * t + 1], q[4 * t + 2],
                         q[4 * t + 3], f, scale),
        "round %d candidate %llu: pool and serial scans differ", round, (unsigned long long)t);

    // the batch's grown candidates as voids, and their pair effects on the pool
    BI.clear();
    place.clear();
    const uint32_t bc = 400;
    std::vector<PairQuery> pq(bc);
    for (uint32_t b = 0; b < bc; b++) {
      PairQuery& c = pq[b];
      c.b = b;
      c.ix = (uint32_t)(rng.next() % n);
      c.iy = (uint32_t)(rng.next() % n);
      c.iz = (uint32_t)(rng.next() % 8);
      c.x = (float)c.ix * l;
      c.y = (float)c.iy * l;
      c.z = (float)c.iz * l;
      c.r = rng.uniform(3.0f, 14.0f);
      c.pole = (int32_t)(scale * 3.0f * (1.0f - f) / l) - 1;
      c.pole_r = (int32_t)(scale * c.r * (1.0f - f) / l) - 1;
      c.need = NEED_NODE | NEED_POLE | NEED_STATE | NEED_POLE_R | NEED_STATE_R;
      BI.add(c.x, c.y, c.z, c.r);
      place.push_back(b);
    }
    constexpr int CAP = 12;
    std::vector<PairEffect> eff(bc * CAP);
    std::vector<int> cnt(bc);
    pool.run(bc, [&](uint64_t a, uint64_t b) {
      for (uint64_t i = a; i < b; i++)
        cnt[i] = pair_effects(BI, place.data(), pq[i], 3.0f, V.max_r, f, scale, l, n,
          &eff[i * CAP], CAP);
    });
    for (uint32_t b = 0; b < bc; b++) {
      PairEffect one[CAP];
      const int c1 = pair_effects(BI, place.data(), pq[b], 3.0f, V.max_r, f, scale, l, n, one, CAP);
      EXPECT(c1 == cnt[b], "round %d candidate %u: %d effects serially, %d on the pool", round,
        b, c1, cnt[b]);
      for (int e = 0; e < c1 && e < cnt[b]; e++)
        EXPECT(one[e].a == eff[b * CAP + e].a && one[e].state == eff[b * CAP + e].state &&
                 one[e].state_r == eff[b * CAP + e].state_r &&
                 one[e].node == eff[b * CAP + e].node,
          "round %d candidate %u effect %d differs", round, b, e);
    }
  }
  // The rung's voids added at once on the pool (add_many) and one by one:
  // the same columns, the same largest radius and the same answers to the
  // commit's probes, over two fillings (the second after clear()).
  RungVoids one, many;
  one.init(box);
  many.init(box);
  for (int fill = 0; fill < 2; fill++) {
    one.clear();
    many.clear();
    const uint32_t nv = fill ? 700 : 2500;
    std::vector<float> vx(nv), vy(nv), vz(nv), vr(nv);
    for (uint32_t i = 0; i < nv; i++) {
      vx[i] = rng.uniform(0, box);
      vy[i] = rng.uniform(0, box);
      vz[i] = i % 4 == 0 ? rng.uniform(0, 3) : rng.uniform(0, box);
      vr[i] = rng.uniform(1.0f, 16.0f);
      one.add(vx[i], vy[i], vz[i], vr[i]);
    }
    many.add_many(
      nv,
      [&](uint64_t i, float* x, float* y, float* z, float* r) {
        *x = vx[i];
        *y = vy[i];
        *z = vz[i];
        *r = vr[i];
      },
      [&](uint64_t m, const std::function<void(uint64_t, uint64_t)>& fn) { pool.run(m, fn); });
    EXPECT(one.cx == many.cx && one.cy == many.cy && one.cz == many.cz && one.r == many.r,
      "fill %d: columns differ", fill);
    EXPECT(one.max_r == many.max_r, "fill %d: largest radius differs", fill);
    const uint64_t m = 4000;
    std::vector<RungProbe> a(m), b(m);
    std::vector<uint32_t> cand(3 * m);
    for (uint64_t t = 0; t < m; t++) {
      cand[3 * t] = (uint32_t)(rng.next() % n);
      cand[3 * t + 1] = (uint32_t)(rng.next() % n);
      cand[3 * t + 2] = t % 3 == 0 ? (uint32_t)(rng.next() % 3) : (uint32_t)(rng.next() % n);
    }
    auto probe = [&](const RungVoids& RV, uint64_t t) {
      const uint32_t ix = cand[3 * t], iy = cand[3 * t + 1], iz = cand[3 * t + 2];
      const float x = (float)ix * l, y = (float)iy * l, z = (float)iz * l;
      OverlapQuery oq;
      overlap_query(oq, x, y, z, 6.0f, RV.max_r, box, f, scale, RV.inv_l);
      return RV.probe(&oq, ix, iy, iz, x, y, z, true, 2, scale, l, n);
    };
    pool.run(m, [&](uint64_t lo, uint64_t hi) {
      for (uint64_t t = lo; t < hi; t++) {
        a[t] = probe(one, t);
        b[t] = probe(many, t);
      }
    });
    // the caller refuses on the first of node, pole, overlap that holds, and
    // the probe stops once that can no longer change: what it decides on
    // must agree (a cell lists its voids in another order in the two lists)
    for (uint64_t t = 0; t < m; t++) {
      bool same = a[t].node == b[t].node;
      if (same && !a[t].node)
        same = a[t].pole == b[t].pole && (a[t].pole || a[t].state == b[t].state);
      EXPECT(same, "fill %d candidate %llu: probes of the two lists differ", fill,
        (unsigned long long)t);
    }
  }

  EXPECT(V.unlinked_reads.load() == 0, "%llu unlinked reads",
    (unsigned long long)V.unlinked_reads.load());

  // a read of a list with voids left unlinked is counted
  Voids W;
  W.init(box);
  W.add(10.0f, 10.0f, 10.0f, 5.0f, 0);
  (void)W.overlap_linked(12.0f, 10.0f, 10.0f, 3.0f, f, scale);
  EXPECT(W.unlinked_reads.load() == 1, "an unlinked read was not counted");

  if (failures) {
    fprintf(stderr, "%d failures\n", failures);
    return 1;
  }
  printf("host harness ok: %d rounds, %u workers, %zu voids\n", rounds, workers, V.cx.size());
  return 0;
}
