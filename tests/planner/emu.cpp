// CPU model of the fused tile pass, for testing the host planner without a GPU.
// TEST INFRASTRUCTURE ONLY (never linked into the product library).
//
// It runs the exact program the planner builds (qsim_fused_plan.hpp: pass
// headers, ops, relayout constants, coefficients) with the semantics of the
// device interpreter in quict_amd/csrc/qsim_fused.hip: per-tile load with the
// load layout, each op on the (thread, register) grid, relayouts through a
// byte-addressed LDS image using the planned slot offsets (every write
// checked for range and collisions), lane factors flushed at relayouts and the
// store, store with the relabeled layout and pending flips.  Single-gate
// passes and the reference use the C restatement in oracle/csrc/cpu_ref_sim.c.
#include "../../quict_amd/csrc/qsim_fused_plan.hpp"

#include <complex>
#include <stdexcept>
#include <string>

extern "C" int cpu_apply_gates(double* vec, int n, const qs_gate* gates, int ngates, int threads);

namespace qs {
static std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

using zc = std::complex<double>;

static int emu_densek_off(uint32_t PM, int KB, int q) {
  int o = 0, k = 0;
  for (int b = 31; b >= 0; --b)
    if ((PM >> b) & 1u) {
      if ((q >> (KB - 1 - k)) & 1) o |= 1 << b;
      ++k;
    }
  return o;
}

static void emu_pass(zc* v, int n, int RB, int TB, uint32_t S, const PassHdr& H, const FOp* P, const double* C,
                     const uint32_t* X) {
  const int NR = 1 << RB, NT = 1 << TB, K = RB + TB;
  const uint64_t tiles = 1ull << (n - K);
  std::vector<zc> a((size_t)NT * NR), f(NT);
  std::vector<zc> lds((size_t)NT * NR);
  std::vector<int> written((size_t)NT * NR);
  std::vector<uint64_t> gin(NT), gout(NT);
  for (int tid = 0; tid < NT; ++tid) {
    uint64_t gi = 0, go = 0;
    for (int b = 0; b < TB; ++b) {
      const uint64_t bit = (tid >> b) & 1u;
      gi |= bit << H.tbits[H.ts0[b]];
      go |= bit << H.tbo[H.ts1[b]];
    }
    gin[tid] = gi;
    gout[tid] = go ^ H.sflip;
  }
  auto cz = [&](int i) { return zc(C[2 * i], C[2 * i + 1]); };
  for (uint64_t t = 0; t < tiles; ++t) {
    const uint64_t tb = insert_zeros(t, H.sorted, K);
    for (int tid = 0; tid < NT; ++tid)
      for (int r = 0; r < NR; ++r) {
        uint64_t g = 0;
        for (int b = 0; b < RB; ++b)
          if ((r >> b) & 1) g |= 1ull << H.tbits[H.rs0[b]];
        a[(size_t)tid * NR + r] = v[tb | gin[tid] | g];
      }
    for (int tid = 0; tid < NT; ++tid) f[tid] = 1.0;
    bool fdirty = false;
    auto flush = [&]() {
      for (int tid = 0; tid < NT; ++tid) {
        for (int r = 0; r < NR; ++r) a[(size_t)tid * NR + r] *= f[tid];
        f[tid] = 1.0;
      }
      fdirty = false;
    };
    for (int o = 0; o < H.nops; ++o) {
      const FOp& op = P[o];
      if (op.kind == F_RELAYOUT) {
        if (fdirty) flush();
        // two rounds through a half-tile image, split on register bit p
        const uint32_t* x = X + op.coef;
        const int p = (int)x[0];
        std::vector<uint32_t> lw(x + 1, x + 1 + RB), tw(x + 1 + RB, x + 1 + RB + TB);
        const uint32_t xf = x[1 + RB + TB];
        const uint32_t* y = x + 2 + RB + TB;
        std::vector<uint32_t> lq(y, y + RB), tq(y + RB, y + RB + TB);
        const uint32_t half = (uint32_t)(NT * NR / 2);
        for (int k = 0; k < 2; ++k) {
          std::fill(written.begin(), written.end(), 0);
          for (int tid = 0; tid < NT; ++tid)
            for (int r = 0; r < NR; ++r) {
              if (((r >> p) & 1) != k) continue;
              uint32_t off = xf;
              for (int b = 0; b < TB; ++b) if ((tid >> b) & 1) off ^= tw[b];
              for (int b = 0; b < RB; ++b) if ((r >> b) & 1) off ^= lw[b];
              if (off % S || off / S >= half) throw std::runtime_error("lds write out of range");
              if (written[off / S]++) throw std::runtime_error("lds write collision");
              lds[off / S] = a[(size_t)tid * NR + r];
            }
          for (int tid = 0; tid < NT; ++tid)
            for (int r = 0; r < NR; ++r) {
              if (((r >> p) & 1) != k) continue;
              uint32_t off = 0;
              for (int b = 0; b < TB; ++b) if ((tid >> b) & 1) off ^= tq[b];
              for (int b = 0; b < RB; ++b) if ((r >> b) & 1) off ^= lq[b];
              if (off % S || off / S >= half) throw std::runtime_error("lds read out of range");
              if (!written[off / S]) throw std::runtime_error("lds read of an unwritten slot");
              a[(size_t)tid * NR + r] = lds[off / S];
            }
        }
        continue;
      }
      const int c0 = op.coef;
      for (int tid = 0; tid < NT; ++tid) {
        zc* A = &a[(size_t)tid * NR];
        const bool tok = (tid & op.cthr) == op.cthrv;
        auto regok = [&](int r) { return ((uint32_t)r & op.creg) == op.cregv; };
        switch (op.kind) {
          case F_DENSE1: {
            const int P0 = op.pos[0];
            if (op.pure) {
              if (op.creg || op.cthr) throw std::runtime_error("real dense1 with controls");
              for (int k = 0; k < 4; ++k)
                if (cz(c0 + k).imag() != 0.0) throw std::runtime_error("real dense1 payload with imaginary part");
            }
            for (int r = 0; r < NR; ++r) {
              if ((r >> P0) & 1 || !regok(r) || !tok) continue;
              const int r1 = r | (1 << P0);
              const zc x0 = A[r], x1 = A[r1];
              A[r] = cz(c0) * x0 + cz(c0 + 1) * x1;
              A[r1] = cz(c0 + 2) * x0 + cz(c0 + 3) * x1;
            }
            break;
          }
          case F_PERM1: {
            const int P0 = op.pos[0];
            for (int r = 0; r < NR; ++r) {
              if ((r >> P0) & 1 || !regok(r) || !tok) continue;
              const int r1 = r | (1 << P0);
              const zc x0 = A[r], x1 = A[r1];
              if (op.pure) {
                A[r] = x1;
                A[r1] = x0;
              } else {
                A[r] = cz(c0) * x1;
                A[r1] = cz(c0 + 1) * x0;
              }
            }
            break;
          }
          case F_DIAG: {
            uint32_t sel = 0;
            for (int k = 0; k < op.ntt; ++k) sel = (sel << 1) | ((tid >> op.tpos[k]) & 1u);
            if (op.pure) {
              f[tid] *= cz(c0 + sel);
              fdirty = true;
            } else {
              for (int r = 0; r < NR; ++r)
                if ((op.creg >> r) & 1u) A[r] *= cz(c0 + sel * NR + r);
            }
            break;
          }
          case F_XX2: {
            const int P0 = op.pos[0], P1 = op.pos[1];
            if (P0 <= P1) throw std::runtime_error("xx2 positions not canonical");
            if (op.creg || op.cthr) throw std::runtime_error("xx2 op with controls");
            for (int r = 0; r < NR; ++r) {
              if (r & ((1 << P0) | (1 << P1))) continue;
              const int idx[4] = {r, r | (1 << P1), r | (1 << P0), r | (1 << P0) | (1 << P1)};
              zc x[4];
              for (int j = 0; j < 4; ++j) x[j] = A[idx[j]];
              for (int s = 0; s < 4; ++s) {
                zc d = cz(c0 + s), o = cz(c0 + 4 + s);
                if (op.pure) {
                  if (d.imag() != 0.0 || o.imag() != 0.0) throw std::runtime_error("xx2 real payload with imaginary part");
                  d = zc(d.real(), 0.0);
                  o = zc(o.real(), 0.0);
                }
                A[idx[s]] = d * x[s] + o * x[s ^ 3];
              }
            }
            break;
          }
          case F_DENSE2:
          case F_SWAP2: {
            const int P0 = op.pos[0], P1 = op.pos[1];
            if (P0 <= P1) throw std::runtime_error("dense2 positions not canonical");
            for (int r = 0; r < NR; ++r) {
              if (r & ((1 << P0) | (1 << P1)) || !regok(r) || !tok) continue;
              const int idx[4] = {r, r | (1 << P1), r | (1 << P0), r | (1 << P0) | (1 << P1)};
              if (op.kind == F_SWAP2) {
                const zc x01 = A[idx[1]], x10 = A[idx[2]];
                A[idx[1]] = op.pure ? x10 : cz(c0) * x10;
                A[idx[2]] = op.pure ? x01 : cz(c0 + 1) * x01;
              } else {
                zc x[4], y[4];
                for (int j = 0; j < 4; ++j) x[j] = A[idx[j]];
                for (int i = 0; i < 4; ++i) {
                  y[i] = 0;
                  for (int j = 0; j < 4; ++j) y[i] += cz(c0 + 4 * i + j) * x[j];
                }
                for (int i = 0; i < 4; ++i) A[idx[i]] = y[i];
              }
            }
            break;
          }
          case F_DENSE3:
          case F_DENSE4: {
            const int KB = op.kind == F_DENSE3 ? 3 : 4, D = 1 << KB;
            uint32_t PM = 0;
            for (int i = 0; i < KB; ++i) PM |= 1u << op.pos[i];
            for (int r = 0; r < NR; ++r) {
              if (((uint32_t)r & PM) || !regok(r) || !tok) continue;
              std::vector<zc> x(D), y(D);
              for (int q = 0; q < D; ++q) x[q] = A[r | emu_densek_off(PM, KB, q)];
              for (int i = 0; i < D; ++i) {
                y[i] = 0;
                for (int j = 0; j < D; ++j) y[i] += cz(c0 + i * D + j) * x[j];
              }
              for (int i = 0; i < D; ++i) A[r | emu_densek_off(PM, KB, i)] = y[i];
            }
            break;
          }
          default:
            throw std::runtime_error("unknown op");
        }
      }
    }
    if (fdirty) flush();
    for (int tid = 0; tid < NT; ++tid)
      for (int r = 0; r < NR; ++r) {
        uint64_t g = 0;
        for (int b = 0; b < RB; ++b)
          if ((r >> b) & 1) g |= 1ull << H.tbo[H.rs1[b]];
        v[(tb ^ gout[tid]) ^ g] = a[(size_t)tid * NR + r];
      }
  }
}
}  // namespace qs

extern "C" const char* emu_last_error() { return qs::g_err.c_str(); }

// Plan and run `gates` on the complex128 vector `vec` (2^n entries) with tile
// shape RB register bits x TB thread bits and L low bits; amp_bytes sets the LDS
// slot offsets (16 for complex128 tiles, 8 for complex64).  vec == NULL: plan
// only.  stats (18 ints): passes, fused passes, ops, ops per kind 0..7, largest
// pass coefficient count, real dense1, lane diagonals, table diagonals,
// lane-controlled ops, diagonal + anti-diagonal two-target ops, passes whose
// store moves qubits between index bits.  QS_EMU_FUSE=2 plans as
// qs_sv_apply_gates with fuse=2 (qubit moves when they save at least 8% of the
// passes); QS_FUSION_REMAP=1 moves qubits whenever that saves a pass.
// QS_EMU_STREAM=1 (plans without moves) runs the passes chunk by chunk as
// the fuse=1 path launches them.
extern "C" int emu_apply(double* vec, int n, const qs_gate* gates, int ngates, int RB, int TB, int L, int amp_bytes,
                         int* stats) {
  using namespace qs;
  try {
    Config c{RB, TB, L};
    if (const char* lr = getenv("QS_FUSION_LOWREGS")) c.LR = atoi(lr);
    if (const char* fu = getenv("QS_EMU_FUSE")) c.remap = atoi(fu) == 2 ? 1 : 0;  // qs_sv_apply_gates' fuse
    if (const char* rm = getenv("QS_FUSION_REMAP")) c.remap = atoi(rm) == 1 ? 2 : 0;
    FusedProgram prog;
    zc* v = reinterpret_cast<zc*>(vec);
    int maxc = 0;
    const char* stream = getenv("QS_EMU_STREAM");
    if (stream && stream[0] == '1' && c.remap == 0) {
      // as qs_sv_apply_gates(fuse=1) runs it: passes run chunk by chunk, from
      // the rebased chunk programs, while the rest is still being planned
      auto chunk = [&](const FusedProgram& sub, int) -> int {
        for (const PassHdr& H : sub.hdrs) {
          maxc = std::max(maxc, H.ncoef);
          if (vec)
            emu_pass(v, n, RB, TB, (uint32_t)amp_bytes, H, sub.ops.data() + H.op0, sub.coef.data() + 2 * H.coef0,
                     sub.aux.data());
        }
        return 0;
      };
      auto single = [&](const qs_gate& g) -> int {
        if (vec) cpu_apply_gates(vec, n, &g, 1, 1);
        return 0;
      };
      int rc = build_streamed(gates, ngates, n, c, (uint32_t)amp_bytes, prog, chunk, single);
      if (rc) return rc;
    } else {
      int rc = build_program(gates, ngates, n, c, (uint32_t)amp_bytes, prog);
      if (rc) return rc;
      for (size_t i = 0; i < prog.plans.size(); ++i) {
        if (!vec) {
          if (prog.plans[i].fused) maxc = std::max(maxc, prog.hdrs[prog.hdr_of[i]].ncoef);
          continue;
        }
        if (!prog.plans[i].fused) {
          cpu_apply_gates(vec, n, &prog.pgates[prog.plans[i].single], 1, 1);
          continue;
        }
        const PassHdr& H = prog.hdrs[prog.hdr_of[i]];
        maxc = std::max(maxc, H.ncoef);
        emu_pass(v, n, RB, TB, (uint32_t)amp_bytes, H, prog.ops.data() + H.op0, prog.coef.data() + 2 * H.coef0,
                 prog.aux.data());
      }
    }
    if (stats) {
      int kinds[9] = {0};
      for (const FOp& o : prog.ops) kinds[o.kind < 9 ? o.kind : 0]++;
      stats[0] = (int)prog.plans.size();
      stats[1] = (int)prog.hdrs.size();
      stats[2] = (int)prog.ops.size();
      for (int k = 0; k < 8; ++k) stats[3 + k] = kinds[k];
      stats[11] = maxc;
      int real1 = 0, ctl1 = 0, lane = 0, dtab = 0, pperm = 0, lanectl = 0;
      for (auto& h : prog.hdrs)
        for (int k = 0; k < h.nops; ++k) {
          const FOp& o = prog.ops[h.op0 + k];
          const double* cc = prog.coef.data() + 2 * (h.coef0 + o.coef);
          if (o.kind == F_DENSE1) {
            bool re = true;
            for (int i = 0; i < 4; ++i) if (cc[2 * i + 1] != 0.0) re = false;
            real1 += re;
            ctl1 += (o.creg || o.cthr);
          }
          if (o.kind == F_DIAG) { lane += o.pure; dtab += !o.pure; }
          if (o.kind == F_PERM1) pperm += o.pure;
          if (o.cthr) lanectl++;
        }
      stats[12] = real1;
      stats[13] = lane;
      stats[14] = dtab;
      stats[15] = lanectl;
      stats[16] = kinds[F_XX2];
      int moved = 0;  // passes whose store moves qubits between index bits
      for (auto& h : prog.hdrs) {
        bool m = false;
        for (int j = 0; j < RB + TB; ++j) m = m || h.tbo[j] != h.tbits[j];
        moved += m;
      }
      stats[17] = moved;
      (void)ctl1;
      (void)pperm;
    }
    return 0;
  } catch (const std::exception& e) {
    fprintf(stderr, "emu error: %s\n", e.what());
    return -99;
  }
}
