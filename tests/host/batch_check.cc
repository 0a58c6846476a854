// Host-side run of the library's batch, graph-capture and solver bookkeeping
// against the stand-in runtime of fake_hip.cc, built with AddressSanitizer
// (tests/test_host_api.py): every host copy, buffer size, stream, event and
// graph the library handles is checked while no kernel runs. The legs follow
// the order of the uneven-batch workload:
//   1. the 1000-field uneven batch: pf_batch_autotune (one graph per branch
//      count 1, 2, 4, 8, 16), graph replays, a timed eager pass;
//   2. one autotuned batch per size class (the fields whose larger side is
//      the class key), created and destroyed one after another;
//   3. plans in the order batch -> CG -> DCT/DST generation -> DFT field.
//
//   batch_check <descriptor file> <n> <class file>
// The descriptor file holds n pf_desc structs; the class file one int32 per
// field (its size class).
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>

#include "parafields_amd.h"

extern "C" long long fake_hip_launches();
extern "C" size_t fake_hip_live_allocations();
extern "C" int hipMalloc(void** p, size_t n);
extern "C" int hipFree(void* p);

static void ok(int st, const char* what)
{
  if (st != PF_OK && st != PF_ENEG) {
    std::printf("FAIL %s: %d %s\n", what, st, pf_last_error());
    std::exit(1);
  }
}

static void run_batch(const std::vector<pf_desc>& d, const char* tag, bool timing)
{
  void* b = nullptr;
  ok(pf_batch_create(d.data(), (int)d.size(), &b), "pf_batch_create");
  std::vector<int64_t> off(d.size() + 1);
  ok(pf_batch_offsets(b, off.data()), "pf_batch_offsets");
  const size_t es = d[0].precision == PF_FP32 ? 4 : 8;
  void* out = nullptr;
  hipMalloc(&out, (size_t)off.back() * es);
  int best = 0;
  double ms = 0;
  ok(pf_batch_autotune(b, 11, 0, PF_TRANSFORM_NONE, 0., out, PF_MEM_DEVICE, &best, &ms),
     "pf_batch_autotune");
  std::vector<pf_eig_stats> st(d.size());
  for (int r = 0; r < 2; r++)
    ok(pf_batch_generate(b, 11, 0, PF_TRANSFORM_NONE, 0., out, PF_MEM_DEVICE, 1, st.data()),
       "graph replay");
  // a new key (seed) re-captures the graph
  ok(pf_batch_generate(b, 12, 0, PF_TRANSFORM_LOGNORMAL, 0., out, PF_MEM_DEVICE, 1, st.data()),
     "re-capture");
  if (timing) {
    ok(pf_batch_set_field_timing(b, 1), "field timing on");
    ok(pf_batch_generate(b, 11, 0, PF_TRANSFORM_NONE, 0., out, PF_MEM_DEVICE, 0, st.data()),
       "timed eager pass");
    std::vector<double> t(d.size());
    ok(pf_batch_field_times(b, t.data()), "pf_batch_field_times");
    ok(pf_batch_set_field_timing(b, 0), "field timing off");
  }
  // host output: staged through the batch's own buffer
  std::vector<float> host((size_t)off.back() * es / sizeof(float));
  ok(pf_batch_generate(b, 11, 0, PF_TRANSFORM_NONE, 0., host.data(), PF_MEM_HOST, 1, st.data()),
     "host output");
  ok(pf_batch_destroy(b), "pf_batch_destroy");
  hipFree(out);
  std::printf("%-28s %5zu fields, best %2d branches OK\n", tag, d.size(), best);
  std::fflush(stdout);
}

static void run_plan(pf_desc d, const char* tag, int batch, bool cg)
{
  void* p = nullptr;
  ok(pf_plan_create(&d, &p), "pf_plan_create");
  pf_sizes s{};
  ok(pf_plan_sizes(p, &s), "pf_plan_sizes");
  pf_eig_stats st{};
  ok(pf_eigs(p, &st), "pf_eigs");
  const size_t es = d.precision == PF_FP32 ? 4 : 8;
  void* out = nullptr;
  hipMalloc(&out, (size_t)batch * s.domain_size * es);
  ok(pf_generate(p, nullptr, 7, 1, batch, PF_TRANSFORM_SIGN, 0.5, out, PF_MEM_DEVICE),
     "pf_generate");
  std::vector<double> noise((size_t)s.noise_per_field * batch, 0.5);
  std::vector<char> host((size_t)batch * s.domain_size * es);
  ok(pf_generate(p, noise.data(), 0, 0, batch, PF_TRANSFORM_NONE, 0., host.data(), PF_MEM_HOST),
     "pf_generate host noise");
  if (cg) {
    std::vector<double> bvec(s.domain_size, 1.0), x(s.domain_size);
    pf_cg_opts o{};
    pf_cg_stats cs{};
    for (int mode : { PF_CG_PLAIN, PF_CG_EMBED_PCG }) {
      o.mode = mode;
      o.max_iterations = 200;
      o.rel_tol = 1e-8;
      ok(pf_cg(p, bvec.data(), x.data(), PF_MEM_HOST, &o, &cs), "pf_cg");
    }
    ok(pf_matvec(p, PF_OP_MULTIPLY, bvec.data(), x.data(), 1, PF_MEM_HOST), "pf_matvec");
  }
  ok(pf_plan_destroy(p), "pf_plan_destroy");
  hipFree(out);
  std::printf("%-28s OK\n", tag);
  std::fflush(stdout);
}

int main(int argc, char** argv)
{
  if (argc < 4)
    return 2;
  const int n = std::atoi(argv[2]);
  std::vector<pf_desc> d(n);
  std::vector<int> cls(n);
  FILE* f = std::fopen(argv[1], "rb");
  if (!f || std::fread(d.data(), sizeof(pf_desc), n, f) != (size_t)n)
    return 2;
  std::fclose(f);
  f = std::fopen(argv[3], "rb");
  if (!f || std::fread(cls.data(), sizeof(int), n, f) != (size_t)n)
    return 2;
  std::fclose(f);
  const size_t live0 = fake_hip_live_allocations();
  run_batch(d, "uneven batch", true);
  std::map<int, std::vector<pf_desc>> by;
  for (int i = 0; i < n; i++)
    by[cls[i]].push_back(d[i]);
  for (auto& kv : by) {
    char tag[64];
    std::snprintf(tag, sizeof tag, "size class %d", kv.first);
    run_batch(kv.second, tag, true);
  }
  // plans after the batches: CG at 64^3 (exponential + nugget), DCT/DST at
  // 128^3, a 1-D DFT field of 2^20 cells
  pf_desc c = d[0];
  c.dim = 3;
  c.cells[0] = c.cells[1] = c.cells[2] = 64;
  c.extensions[0] = c.extensions[1] = c.extensions[2] = 1.;
  c.covariance = PF_COV_EXPONENTIAL;
  c.anisotropy = PF_ANISO_NONE;
  c.corr_length[0] = 0.1;
  c.matrix_backend = PF_MATRIX_DCT;
  c.field_backend = PF_FIELD_R2C;
  c.transposed = -1;
  c.precision = PF_FP64;
  c.nugget = 1e-3;
  c.r2c_check_all = 0;
  run_plan(c, "CG 64^3", 2, true);
  pf_desc t = c;
  t.cells[0] = t.cells[1] = t.cells[2] = 128;
  t.field_backend = PF_FIELD_DCTDST;
  t.precision = PF_FP32;
  t.nugget = 0.;
  run_plan(t, "DCT/DST 128^3", 3, false);
  pf_desc o = c;
  o.dim = 1;
  o.cells[0] = 1 << 20;
  o.cells[1] = o.cells[2] = 1;
  o.matrix_backend = PF_MATRIX_DFT;
  o.field_backend = PF_FIELD_DFT;
  run_plan(o, "1-D DFT 2^20", 3, true);
  o.precision = PF_FP32;
  run_plan(o, "1-D DFT 2^20 FP32", 5, false);
  std::printf("launches %lld, device allocations still live %zu (before %zu)\n",
              fake_hip_launches(), fake_hip_live_allocations(), live0);
  std::printf("all OK\n");
  return 0;
}
