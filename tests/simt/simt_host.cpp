// CPU emulation harness of the perturbation kernels' per-item device code.
//
// Compiles the device headers (disco-eb_amd/csrc/deb_forward.h,
// deb_adjoint.h) for the host against tests/simt/include/hip/hip_runtime.h:
// one wavefront = 64 host threads; cross-lane operations go through a
// per-wave exchange buffer.  Built with AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/simt/Makefile), it checks the device code
// for out-of-bounds accesses, uninitialised-pointer use and undefined
// arithmetic, and its results are compared with the oracle by
// tests/test_simt_host.py.  Test infrastructure only.
//
//   simt_host <in.bin> <out.bin>
//
// in.bin (little-endian, written by tests/test_simt_host.py):
//   int32 mode (0: forward solve, 1: taped forward + reverse sweep, 2: the
//   row coefficients' wave form against the row switch)
//   deb_config_t cfg; deb_cosmo_t cosmo; int64 stride; double tables[stride];
//   int32 desc[14]; int32 nq; double q[max(nq,1)], w[max(nq,1)];
//   int32 nk; double k[nk]; int32 nout; double tau[nout];
//   mode 1 adds: int32 ckpt; int32 ndir; double dcos[ndir][24];
//                double dtab[ndir][stride]; double dtau[ndir][nout];
//                double ybar[nk][nout][20]
// out.bin: int32 status[nk]; int32 nsteps[nk]; double y[nk][nout][width];
//   mode 1 adds: double grad[nk][ndir]; mode 2 writes only int32 mismatches,
//   int32 rows checked
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../disco-eb_amd/csrc/deb_adjoint.h"
#include "../../disco-eb_amd/csrc/deb_host.h"

namespace deb {
thread_local std::string g_err;
}

using namespace deb;

namespace {

struct Reader {
  FILE *f;
  template <class T>
  void get(T *p, size_t n) {
    if (n && fread(p, sizeof(T), n, f) != n) {
      fprintf(stderr, "simt_host: short input\n");
      exit(2);
    }
  }
  template <class T>
  T one() {
    T v;
    get(&v, 1);
    return v;
  }
};

// runs body(lane) on 64 threads sharing one wave's exchange buffer
template <class F>
void run_wave(F body) {
  simt::Wave w;
  std::vector<std::thread> th;
  for (int l = 0; l < WAVE; ++l)
    th.emplace_back([&, l] {
      simt::lane = l;
      simt::parity = 0;
      simt::wave = &w;
      body(l);
    });
  for (auto &t : th) t.join();
}

template <int NS>
int run(const deb_config_t &cfg, const Layout &L, Reader &in, FILE *out, int mode, const deb_cosmo_t &cosmo,
        const std::vector<double> &tables, int64_t stride, const int32_t *desc, const std::vector<double> &q,
        const std::vector<double> &w, const std::vector<double> &k, const std::vector<double> &tau) {
  const int nk = (int)k.size(), nout = (int)tau.size();
  const int n = 12 + cfg.lmaxg + cfg.lmaxgp + cfg.lmaxr + cfg.nqmax * (cfg.lmaxnu + 1);
  const int width = cfg.return_full ? n : DEB_NOUT_FIELDS;
  std::vector<double> y((size_t)nk * nout * width, -12345.0);
  std::vector<int32_t> status(nk, -9), nsteps(nk, 0);
  Params<NS> P;
  if (int r = fill_params<NS>(P, &cfg, L, 1, &cosmo, tables.data(), stride, desc, q.data(), w.data(), nk, k.data(),
                              nout, tau.data()))
    return r;
  P.y_out = y.data();
  P.status = status.data();
  P.nsteps = nsteps.data();
  if (mode == 2) {
    // the wave form of the row coefficients against the row switch, bit for
    // bit, for every lane and slot at several a and both dark-energy modes
    int32_t bad = 0, checked = 0;
    for (int ki = 0; ki < nk; ++ki) {
      run_wave([&](int lane) {
        Ctx cx;
        make_ctx<NS>(cx, P, 0, k[ki]);
        uint32_t rd[NS];
        for (int s = 0; s < NS; ++s) rd[s] = P.rowdesc[s][lane];
        int nb = 0, nc = 0;
        for (int lam = 0; lam < 2; ++lam) {
          cx.lambda = lam == 1;
          for (double a = 1e-7; a <= 1.0; a *= 7.0) {
            const Bg<double> b = background<double>(cx, a);
            const double invtau = 1.0 / (30.0 + 1e4 * a);
            dual ad;
            ad.v = a;
            ad.d = 1.0;
            const Bg<dual> bd = background<dual>(cx, ad);
            for (int s = 0; s < NS; ++s) {
              const Coef<double> r0 = row_coef<double>(rd[s], b, cx, invtau);
              const Coef<double> r1 = row_coef_wave<double>(rd[s], b, cx, invtau, s == 0);
              nb += std::memcmp(&r0, &r1, sizeof(r0)) != 0;
              // in dual arithmetic along a: equal values and derivatives
              const Coef<dual> d0 = row_coef<dual>(rd[s], bd, cx, invtau);
              const Coef<dual> d1 = row_coef_wave<dual>(rd[s], bd, cx, invtau, s == 0);
              const dual *p0 = &d0.sub, *p1 = &d1.sub;
              for (int f = 0; f < 11; ++f) nb += !(p0[f].v == p1[f].v && p0[f].d == p1[f].d);
              nc += 2;
            }
          }
        }
        __atomic_add_fetch(&bad, nb, __ATOMIC_RELAXED);
        __atomic_add_fetch(&checked, nc, __ATOMIC_RELAXED);
      });
    }
    fwrite(&bad, 4, 1, out);
    fwrite(&checked, 4, 1, out);
    return 0;
  }
  if (mode == 0) {
    for (int ki = 0; ki < nk; ++ki) {
      std::vector<double> lds(fwd_lds_doubles<NS>(), -777.0), pcr(pcr_lds_doubles<NS>(), -444.0);
      run_wave([&](int) { integrate_mode<NS>(P, 0, ki, lds.data(), pcr.data()); });
    }
    fwrite(status.data(), 4, nk, out);
    fwrite(nsteps.data(), 4, nk, out);
    fwrite(y.data(), 8, y.size(), out);
    return 0;
  }
  const int ckpt = in.one<int32_t>();
  const int ndir = in.one<int32_t>();
  std::vector<double> dcos((size_t)ndir * 24), dtab((size_t)ndir * stride), dtau((size_t)ndir * nout),
      ybar((size_t)nk * nout * DEB_NOUT_FIELDS);
  in.get(dcos.data(), dcos.size());
  in.get(dtab.data(), dtab.size());
  in.get(dtau.data(), dtau.size());
  in.get(ybar.data(), ybar.size());
  const long long n_states = 200000;
  const long long chunks = n_states / ckpt + nk + 1;
  const size_t tape_bytes = (size_t)(tape_pool_off(nk) + chunks * tape_chunk_bytes(NS, ckpt));
  std::vector<char> tape(tape_bytes, 0);
  TapeView tv;
  if (int r = tape_view(tv, tape.data(), tape_bytes, nk, NS, ckpt)) return r;
  for (int ki = 0; ki < nk; ++ki) {
    std::vector<double> lds(fwd_lds_doubles<NS>(), -777.0), pcr(pcr_lds_doubles<NS>(), -444.0);
    run_wave([&](int) { integrate_mode<NS, true>(P, 0, ki, lds.data(), pcr.data(), &tv); });
  }
  AdjArgs A;
  std::memset(&A, 0, sizeof(A));
  A.ndir = ndir;
  A.dcos = dcos.data();
  A.dtab = dtab.data();
  A.dtau = dtau.data();
  A.ybar = ybar.data();
  A.status = status.data();
  std::vector<double> grad((size_t)nk * ndir, -999.0);
  A.grad = grad.data();
  A.tape = tv;
  A.n_in = output_inputs(cfg, A.in_idx);
  for (int ki = 0; ki < nk; ++ki) {
    std::vector<double> ring(ring_doubles(NS, ckpt), -555.0);
    std::vector<double> KB(7 * NS * WAVE, -666.0), pcr(pcr_lds_doubles<NS>(), -444.0);
    Ctx cxs;  // the wave's shared context record (every lane stores the same values)
    std::memset(&cxs, 0xa5, sizeof(cxs));
    // the output conversion's reverse stages its data in the first two blocks
    // of the stage-cotangent buffer, as the kernel does
    run_wave([&](int) {
      adjoint_mode<NS>(P, A, 0, ki, ring.data(), KB.data(), KB.data(), KB.data() + NS * WAVE, pcr.data(), cxs);
    });
  }
  fwrite(status.data(), 4, nk, out);
  fwrite(nsteps.data(), 4, nk, out);
  fwrite(y.data(), 8, y.size(), out);
  fwrite(grad.data(), 8, grad.size(), out);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: simt_host in.bin out.bin\n");
    return 2;
  }
  Reader in{fopen(argv[1], "rb")};
  if (!in.f) return 2;
  const int mode = in.one<int32_t>();
  const deb_config_t cfg = in.one<deb_config_t>();
  const deb_cosmo_t cosmo = in.one<deb_cosmo_t>();
  const int64_t stride = in.one<int64_t>();
  std::vector<double> tables(stride);
  in.get(tables.data(), tables.size());
  int32_t desc[14];
  in.get(desc, 14);
  const int nq = in.one<int32_t>();
  std::vector<double> q(std::max(nq, 1)), w(std::max(nq, 1));
  in.get(q.data(), q.size());
  in.get(w.data(), w.size());
  const int nk = in.one<int32_t>();
  std::vector<double> k(nk);
  in.get(k.data(), nk);
  const int nout = in.one<int32_t>();
  std::vector<double> tau(nout);
  in.get(tau.data(), nout);
  Layout L;
  if (build_layout(cfg, L)) {
    fprintf(stderr, "simt_host: %s\n", g_err.c_str());
    return 2;
  }
  FILE *out = fopen(argv[2], "wb");
  if (!out) return 2;
  int r = -1;
  switch (L.ns) {
    case 1: r = run<1>(cfg, L, in, out, mode, cosmo, tables, stride, desc, q, w, k, tau); break;
    case 2: r = run<2>(cfg, L, in, out, mode, cosmo, tables, stride, desc, q, w, k, tau); break;
    case 4: r = run<4>(cfg, L, in, out, mode, cosmo, tables, stride, desc, q, w, k, tau); break;
    case 5: r = run<5>(cfg, L, in, out, mode, cosmo, tables, stride, desc, q, w, k, tau); break;
    default: fprintf(stderr, "simt_host: slot count %d not built\n", L.ns); return 2;
  }
  fclose(out);
  if (r) fprintf(stderr, "simt_host: %s\n", g_err.c_str());
  return r ? 3 : 0;
}
