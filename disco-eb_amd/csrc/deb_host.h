// Host-side helpers shared by the kernel translation units: lane layout of
// the state, parameter blocks, argument checks and error reporting.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "deb_device.h"

namespace deb {

// ---------------------------------------------------------------------------
// host side: lane layout, argument checks, launches

extern thread_local std::string g_err;

inline int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

struct Layout {
  int ns = 0;
  std::vector<uint32_t> rd;  // [ns][64]
};

inline int build_layout(const deb_config_t &c, Layout &L) {
  const int lg = c.lmaxg, lgp = c.lmaxgp, lr = c.lmaxr, lnu = c.lmaxnu, nq = c.nqmax;
  if (lg < 3 || lgp < 3 || lr < 3 || (nq > 0 && lnu < 3))
    return fail(DEB_E_UNSUPPORTED, "hierarchy truncations must be >= 3");
  if (nq < 0 || nq > DEB_MAXNQ) return fail(DEB_E_UNSUPPORTED, "nqmax outside [0, 16]");
  if (8 + lg + 2 > WAVE) return fail(DEB_E_UNSUPPORTED, "lmaxg > 54 does not fit one wavefront");
  if (lgp + 1 > WAVE || lr + 1 > WAVE || lnu + 1 > WAVE)
    return fail(DEB_E_UNSUPPORTED, "a hierarchy longer than 64 multipoles does not fit one wavefront");
  const int n = 12 + lg + lgp + lr + nq * (lnu + 1);
  const int ig = 7, igp = 8 + lg, ir = 9 + lg + lgp, iq0 = 10 + lg + lgp + lr;
  struct Chain {
    std::vector<uint32_t> rows;
  };
  std::vector<Chain> chains;
  {  // polarization
    Chain ch;
    for (int l = 0; l <= lgp; ++l) {
      uint32_t t = l == 0 ? R_G0 : (l == 2 ? R_G2 : (l == lgp ? R_GLMAX : R_GL));
      ch.rows.push_back(rd_make(t, l, 0, igp + l));
    }
    chains.push_back(ch);
  }
  {  // massless neutrinos
    Chain ch;
    for (int l = 0; l <= lr; ++l) {
      uint32_t t = l == 0 ? R_DR : (l == 1 ? R_TR : (l == 2 ? R_FR2 : (l == lr ? R_FRLMAX : R_FRL)));
      ch.rows.push_back(rd_make(t, l, 0, ir + l));
    }
    chains.push_back(ch);
  }
  for (int i = 0; i < nq; ++i) {  // massive neutrinos, one chain per momentum bin
    Chain ch;
    for (int l = 0; l <= lnu; ++l) {
      uint32_t t = l == 0 ? R_Q0 : (l == 1 ? R_Q1 : (l == 2 ? R_Q2 : (l == lnu ? R_QLMAX : R_QL)));
      ch.rows.push_back(rd_make(t, l, i, iq0 + l * nq + i));
    }
    chains.push_back(ch);
  }
  std::vector<std::vector<uint32_t>> slots(1, std::vector<uint32_t>(WAVE, rd_make(R_NONE, 0, 0, 0)));
  auto &s0 = slots[0];
  s0[L_A] = rd_make(R_A, 0, 0, 0);
  s0[L_AHP] = rd_make(R_AHP, 0, 0, 1);
  s0[L_ETA] = rd_make(R_ETA, 0, 0, 2);
  s0[L_DG] = rd_make(R_DG, 0, 0, ig);
  s0[L_DC] = rd_make(R_DC, 0, 0, 3);
  s0[L_TC] = rd_make(R_TC, 0, 0, 4);
  s0[L_DQ] = rd_make(R_DQ, 0, 0, n - 2);
  s0[L_TQ] = rd_make(R_TQ, 0, 0, n - 1);
  s0[L_DB] = rd_make(R_DB, 0, 0, 5);
  s0[L_TB] = rd_make(R_TB, 1, 0, 6);
  s0[L_TG] = rd_make(R_TG, 1, 0, ig + 1);
  for (int l = 2; l <= lg; ++l) {
    uint32_t t = l == 2 ? R_F2 : (l == lg ? R_FLMAX : R_FL);
    s0[L_TG + l - 1] = rd_make(t, l, 0, ig + l);
  }
  std::vector<int> used(1, L_TG + lg);
  std::vector<int> order(chains.size());
  for (size_t i = 0; i < chains.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return chains[a].rows.size() > chains[b].rows.size(); });
  for (int ci : order) {
    const int len = (int)chains[ci].rows.size();
    int slot = -1;
    for (size_t s = 0; s < slots.size(); ++s)
      if (used[s] + len <= WAVE) {
        slot = (int)s;
        break;
      }
    if (slot < 0) {
      slots.emplace_back(WAVE, rd_make(R_NONE, 0, 0, 0));
      used.push_back(0);
      slot = (int)slots.size() - 1;
    }
    for (int j = 0; j < len; ++j) slots[slot][used[slot] + j] = chains[ci].rows[j];
    used[slot] += len;
  }
  if ((int)slots.size() > MAXNS) return fail(DEB_E_UNSUPPORTED, "state does not fit 6 lane slots");
  L.ns = (int)slots.size();
  L.rd.clear();
  for (auto &s : slots) L.rd.insert(L.rd.end(), s.begin(), s.end());
  return 0;
}

template <int NS>
int fill_params(Params<NS> &P, const deb_config_t *cfg, const Layout &L, int n_cosmo, const deb_cosmo_t *cosmo,
                const double *tables, int64_t stride, const int32_t *desc, const double *q, const double *w, int nk,
                const double *kmodes, int nout, const double *tau_out) {
  std::memset(&P, 0, sizeof(P));
  P.cfg = *cfg;
  P.nq = cfg->nqmax;
  P.n = 12 + cfg->lmaxg + cfg->lmaxgp + cfg->lmaxr + cfg->nqmax * (cfg->lmaxnu + 1);
  P.n_cosmo = n_cosmo;
  P.nk = nk;
  P.nout = nout;
  P.table_stride = stride;
  for (int s = 0; s < DEB_NSPLINE; ++s) {
    P.desc[s][0] = desc[2 * s];
    P.desc[s][1] = desc[2 * s + 1];
    if (desc[2 * s + 1] < 2 || desc[2 * s] < 0 || desc[2 * s] + 3 * (int64_t)desc[2 * s + 1] > stride)
      return fail(DEB_E_ARG, "spline descriptor outside the table stride");
  }
  for (int i = 0; i < P.nq; ++i) {
    P.q[i] = q[i];
    P.w[i] = w[i];
    P.dlnf[i] = -q[i] / (1.0 + std::exp(-q[i]));
  }
  P.cosmo = cosmo;
  P.tables = tables;
  P.kmodes = kmodes;
  P.tau_out = tau_out;
  for (int s = 0; s < NS; ++s)
    for (int l = 0; l < WAVE; ++l) P.rowdesc[s][l] = L.rd[(size_t)s * WAVE + l];
  return 0;
}

inline int check_common(const deb_config_t *cfg, int n_cosmo, const deb_cosmo_t *cosmo, const double *tables,
                 const int32_t *desc, const double *q, const double *w) {
  if (!cfg || !cosmo || !tables || !desc) return fail(DEB_E_ARG, "null argument");
  if (n_cosmo < 1) return fail(DEB_E_ARG, "n_cosmo must be >= 1");
  if (cfg->nqmax > 0 && (!q || !w)) return fail(DEB_E_ARG, "q/w required when nqmax > 0");
  if (cfg->solver != DEB_SOLVER_RODAS5) return fail(DEB_E_ARG, "unknown solver");
  if (!(cfg->rtol > 0) || !(cfg->atol >= 0)) return fail(DEB_E_ARG, "rtol must be > 0 and atol >= 0");
  if (cfg->max_steps < 1) return fail(DEB_E_ARG, "max_steps must be >= 1");
  return 0;
}

inline int hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) return fail(DEB_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

inline int tape_view(TapeView &tv, void *tape, size_t bytes, long long n_modes, int ns, int ckpt) {
  if (ckpt < 1 || ckpt > TAPE_MAX_CKPT) return fail(DEB_E_ARG, "checkpoint interval outside [1, 4096]");
  const long long pool = tape_pool_off(n_modes);
  const long long cb = tape_chunk_bytes(ns, ckpt);
  if (!tape || (long long)bytes < pool + cb) return fail(DEB_E_ARG, "tape smaller than deb_tape_bytes");
  char *base = static_cast<char *>(tape);
  tv.counter = reinterpret_cast<unsigned long long *>(base);
  tv.capacity = ((long long)bytes - pool) / cb;
  tv.last = reinterpret_cast<long long *>(base + tape_last_off());
  tv.count = reinterpret_cast<int32_t *>(base + tape_count_off(n_modes));
  tv.pool = base + pool;
  tv.chunk_bytes = cb;
  tv.state_off = tape_state_off(ckpt);
  tv.ckpt = ckpt;
  return 0;
}

// DEB_NS_MASK (bit NS) limits which slot counts are compiled (development
// builds); the default compiles all of 1..6.
#ifndef DEB_NS_MASK
#define DEB_NS_MASK 0x7e
#endif
#define DEB_NS_CASE(V, CALL)                                          \
  case V: {                                                           \
    if constexpr ((DEB_NS_MASK >> V) & 1) {                           \
      constexpr int NS = V;                                           \
      return CALL;                                                    \
    } else {                                                          \
      return fail(DEB_E_UNSUPPORTED, "slot count not compiled in"); \
    }                                                                 \
  }
#define DEB_DISPATCH(NSV, CALL)                                      \
  switch (NSV) {                                                     \
    DEB_NS_CASE(1, CALL)                                             \
    DEB_NS_CASE(2, CALL)                                             \
    DEB_NS_CASE(3, CALL)                                             \
    DEB_NS_CASE(4, CALL)                                             \
    DEB_NS_CASE(5, CALL)                                             \
    DEB_NS_CASE(6, CALL)                                             \
    default: return fail(DEB_E_UNSUPPORTED, "unsupported slot count"); \
  }


}  // namespace deb
