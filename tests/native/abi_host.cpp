// A native host program that drives the library through its C ABI only
// (include/redzone.h), with HIP for device memory, the way a C++ caller of
// the reference's remove(data, n, removeList, k) would.  It checks results
// with the invariants the reference's own test harness checks
// (DataStruct.hxx:154-191: length, the kept items, the in-place property)
// on the reference's test list A[i] = i (DataStruct.hxx:77).
//
// Exit codes: 0 all checks passed, 1 a check failed, 77 no usable GPU.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <algorithm>
#include <numeric>
#include <random>
#include <vector>

#include "redzone.h"

static int failures = 0;
#define CHECK(cond, ...)                          \
  do {                                            \
    if (!(cond)) {                                \
      fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      fprintf(stderr, __VA_ARGS__);               \
      fprintf(stderr, "\n");                      \
      ++failures;                                 \
    }                                             \
  } while (0)
#define HIPCHECK(x)                                                              \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) {                                                      \
      fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
      return 1;                                                                  \
    }                                                                            \
  } while (0)

// k distinct removals in [0, n): a partial Fisher-Yates shuffle, as the
// reference's generator draws them
static std::vector<int32_t> removals(int64_t n, int64_t k, uint64_t seed) {
  std::vector<int32_t> all(n);
  std::iota(all.begin(), all.end(), 0);
  std::mt19937_64 rng(seed);
  for (int64_t i = 0; i < k; ++i) {
    std::uniform_int_distribution<int64_t> pick(i, n - 1);
    std::swap(all[i], all[pick(rng)]);
  }
  all.resize(k);
  return all;
}

// the Redzone contract on A[i] = i: the first n - k items are the kept
// indices (each once), and every kept index below n - k stayed in place
static void check_redzone(const std::vector<uint32_t>& out, int64_t n, const std::vector<int32_t>& R,
                          const char* what) {
  const int64_t k = (int64_t)R.size();
  std::vector<char> removed(n, 0), seen(n, 0);
  for (int32_t r : R) removed[r] = 1;
  int64_t bad = 0;
  for (int64_t i = 0; i < n - k; ++i) {
    const uint32_t v = out[i];
    if (v >= (uint64_t)n || removed[v] || seen[v]) { ++bad; continue; }
    seen[v] = 1;
    if (!removed[i] && v != (uint32_t)i) ++bad;  // in place
  }
  CHECK(bad == 0, "%s: %lld items break the contract (n=%lld k=%lld)", what, (long long)bad, (long long)n,
        (long long)k);
}

int main() {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    rz_ctx* ctx = nullptr;
    const rz_status st = rz_ctx_create(&ctx, 0);
    printf("no GPU: rz_ctx_create -> %s\n", rz_status_string(st));
    return st != RZ_OK && ctx == nullptr ? 77 : 1;
  }
  rz_ctx* ctx = nullptr;
  CHECK(rz_ctx_create(&ctx, 0) == RZ_OK, "rz_ctx_create");
  if (!ctx) return 1;
  hipStream_t s;
  HIPCHECK(hipStreamCreate(&s));

  // 1. rz_remove_u32 on the reference's config-1 size and on a 2^24 list
  for (auto nk : {std::pair<int64_t, int64_t>{1 << 20, 1 << 14}, {1 << 24, 1 << 18}, {100000, 99999}}) {
    const int64_t n = nk.first, k = nk.second;
    std::vector<uint32_t> A(n);
    std::iota(A.begin(), A.end(), 0u);
    const std::vector<int32_t> R = removals(n, k, 2024 + n);
    uint32_t* dA;
    int32_t* dR;
    HIPCHECK(hipMalloc(&dA, n * 4));
    HIPCHECK(hipMalloc(&dR, k * 4));
    HIPCHECK(hipMemcpy(dA, A.data(), n * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dR, R.data(), k * 4, hipMemcpyHostToDevice));
    CHECK(rz_remove_u32(ctx, dA, n, dR, k, 0, s) == RZ_OK, "rz_remove_u32 n=%lld", (long long)n);
    uint32_t bits = 1;
    CHECK(rz_check_errors(ctx, s, &bits) == RZ_OK && bits == 0, "device error bits %#x", bits);
    std::vector<uint32_t> out(n);
    HIPCHECK(hipMemcpy(out.data(), dA, n * 4, hipMemcpyDeviceToHost));
    check_redzone(out, n, R, "rz_remove_u32");

    // 1b. the allocation-free entry point on the same list: no context, a
    // caller-owned error word
    uint32_t* dErr;
    HIPCHECK(hipMalloc(&dErr, 4));
    HIPCHECK(hipMemset(dErr, 0, 4));
    HIPCHECK(hipMemcpy(dA, A.data(), n * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dR, R.data(), k * 4, hipMemcpyHostToDevice));
    CHECK(rz_remove_noalloc(dA, n, 4, dR, k, 0, dErr, s) == RZ_OK, "rz_remove_noalloc n=%lld", (long long)n);
    HIPCHECK(hipStreamSynchronize(s));
    uint32_t err = 1;
    HIPCHECK(hipMemcpy(&err, dErr, 4, hipMemcpyDeviceToHost));
    CHECK(err == 0, "rz_remove_noalloc error bits %#x", err);
    HIPCHECK(hipMemcpy(out.data(), dA, n * 4, hipMemcpyDeviceToHost));
    check_redzone(out, n, R, "rz_remove_noalloc");
    HIPCHECK(hipFree(dErr));

    // 1c. sizes in device memory (the reference's TestGPUWide(A, sizeA, R, sizeR)
    // shape): n and k written to the device, the call sized for them
    int32_t* dNK;
    HIPCHECK(hipMalloc(&dNK, 8));
    const int32_t sizes[2] = {(int32_t)n, (int32_t)k};
    HIPCHECK(hipMemcpy(dNK, sizes, 8, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dA, A.data(), n * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dR, R.data(), k * 4, hipMemcpyHostToDevice));
    CHECK(rz_remove_sized(ctx, dA, n, 4, dR, k, dNK, dNK + 1, 0, s) == RZ_OK, "rz_remove_sized n=%lld",
          (long long)n);
    bits = 1;
    CHECK(rz_check_errors(ctx, s, &bits) == RZ_OK && bits == 0, "rz_remove_sized error bits %#x", bits);
    HIPCHECK(hipMemcpy(out.data(), dA, n * 4, hipMemcpyDeviceToHost));
    check_redzone(out, n, R, "rz_remove_sized");
    HIPCHECK(hipFree(dNK));

    // 2. the O(n) partition on the same list: stable, so the survivors are the
    // kept indices in ascending order
    HIPCHECK(hipMemcpy(dA, A.data(), n * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dR, R.data(), k * 4, hipMemcpyHostToDevice));
    CHECK(rz_partition_on(ctx, dA, n, 4, dR, k, 0, s) == RZ_OK, "rz_partition_on");
    HIPCHECK(hipStreamSynchronize(s));
    HIPCHECK(hipMemcpy(out.data(), dA, n * 4, hipMemcpyDeviceToHost));
    std::vector<char> removed(n, 0);
    for (int32_t r : R) removed[r] = 1;
    int64_t j = 0, bad = 0;
    for (int64_t i = 0; i < n; ++i)
      if (!removed[i] && out[j++] != (uint32_t)i) ++bad;
    CHECK(bad == 0 && j == n - k, "rz_partition_on: %lld items out of order", (long long)bad);
    HIPCHECK(hipFree(dA));
    HIPCHECK(hipFree(dR));
  }

  // 3. three lists in one persistent launch
  {
    const int64_t ns[3] = {5000, 1 << 16, 300000}, ks[3] = {17, 1 << 12, 150000};
    std::vector<std::vector<int32_t>> Rs;
    rz_list_desc descs[3];
    std::vector<uint32_t*> dAs;
    std::vector<int32_t*> dRs;
    for (int l = 0; l < 3; ++l) {
      std::vector<uint32_t> A(ns[l]);
      std::iota(A.begin(), A.end(), 0u);
      Rs.push_back(removals(ns[l], ks[l], 77 + l));
      uint32_t* dA;
      int32_t* dR;
      HIPCHECK(hipMalloc(&dA, ns[l] * 4));
      HIPCHECK(hipMalloc(&dR, ks[l] * 4));
      HIPCHECK(hipMemcpy(dA, A.data(), ns[l] * 4, hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(dR, Rs[l].data(), ks[l] * 4, hipMemcpyHostToDevice));
      descs[l] = rz_list_desc{dA, dR, ns[l], ks[l], 4u, 0u};
      dAs.push_back(dA);
      dRs.push_back(dR);
    }
    CHECK(rz_remove_batched(ctx, descs, 3, 0, s) == RZ_OK, "rz_remove_batched");
    HIPCHECK(hipStreamSynchronize(s));
    for (int l = 0; l < 3; ++l) {
      std::vector<uint32_t> out(ns[l]);
      HIPCHECK(hipMemcpy(out.data(), dAs[l], ns[l] * 4, hipMemcpyDeviceToHost));
      check_redzone(out, ns[l], Rs[l], "rz_remove_batched");
    }

    // 3b. the same lists through rz_remove_auto_batched with a threshold of
    // 2^12 removals: list 0 in the batch, lists 1 and 2 on their own launches
    // with the k < n/2 rule (no network is installed in this context), so
    // list 1 runs Redzone and list 2 (2k = n) the stable partition
    for (int l = 0; l < 3; ++l) {
      std::vector<uint32_t> A(ns[l]);
      std::iota(A.begin(), A.end(), 0u);
      HIPCHECK(hipMemcpy(dAs[l], A.data(), ns[l] * 4, hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(dRs[l], Rs[l].data(), ks[l] * 4, hipMemcpyHostToDevice));
    }
    int32_t method[3] = {-1, -1, -1};
    CHECK(rz_remove_auto_batched(ctx, nullptr, descs, 3, 1 << 12, 0, method, s) == RZ_OK, "rz_remove_auto_batched");
    HIPCHECK(hipStreamSynchronize(s));
    CHECK(method[0] == RZ_METHOD_REDZONE && method[1] == RZ_METHOD_REDZONE && method[2] == RZ_METHOD_PARTITION,
          "rz_remove_auto_batched methods %d %d %d", method[0], method[1], method[2]);
    for (int l = 0; l < 3; ++l) {
      std::vector<uint32_t> out(ns[l]);
      HIPCHECK(hipMemcpy(out.data(), dAs[l], ns[l] * 4, hipMemcpyDeviceToHost));
      if (method[l] == RZ_METHOD_PARTITION) {
        std::vector<char> removed(ns[l], 0);
        for (int32_t r : Rs[l]) removed[r] = 1;
        int64_t j = 0, bad = 0;
        for (int64_t i = 0; i < ns[l]; ++i)
          if (!removed[i] && out[j++] != (uint32_t)i) ++bad;
        CHECK(bad == 0 && j == ns[l] - ks[l], "rz_remove_auto_batched partition: %lld out of order", (long long)bad);
      } else {
        check_redzone(out, ns[l], Rs[l], "rz_remove_auto_batched");
      }
      HIPCHECK(hipFree(dAs[l]));
      HIPCHECK(hipFree(dRs[l]));
    }
  }

  // 4. the error convention: status codes, nothing touched
  {
    uint32_t* dA;
    int32_t* dR;
    HIPCHECK(hipMalloc(&dA, 64 * 4));
    HIPCHECK(hipMalloc(&dR, 4 * 4));
    const int32_t dup[4] = {3, 9, 3, 1};
    HIPCHECK(hipMemcpy(dR, dup, sizeof(dup), hipMemcpyHostToDevice));
    CHECK(rz_remove_u32(ctx, dA, 3, dR, 4, 0, s) == RZ_EINVAL, "k > n must be RZ_EINVAL");
    CHECK(rz_remove_u32(ctx, dA, 64, dR, 4, RZ_VALIDATE, s) == RZ_EDUP, "a duplicate must be RZ_EDUP");
    CHECK(rz_remove_noalloc(dA, 1 << 23, 4, dR, RZ_NOALLOC_MAX_K + 1, 0, nullptr, s) == RZ_ETOOBIG,
          "k above the fixed scratch must be RZ_ETOOBIG");
    CHECK(rz_remove_sized(ctx, dA, 64, 4, dR, 4, nullptr, dR, 0, s) == RZ_EINVAL, "rz_remove_sized needs d_n");
    CHECK(rz_remove_sized(ctx, dA, 64, 4, dR, 4, dR, dR, RZ_VALIDATE, s) == RZ_EINVAL,
          "rz_remove_sized cannot validate (k is on the device)");
    // the batched method choice: no network is installed in this context, so the k < n/2 rule
    int32_t choice[4] = {-1, -1, -1, -1};
    CHECK(rz_dispatch_collect(ctx, choice, 4) == RZ_EINVAL, "rz_dispatch_collect with nothing issued");
    const rz_list_desc bad[1] = {{dA, dR, 3, 4, 4, 0}};
    CHECK(rz_dispatch_choose(ctx, nullptr, bad, 1, s) == RZ_EINVAL, "rz_dispatch_choose with k > n");
    const rz_list_desc lists[4] = {{dA, dR, 64, 4, 4, 0}, {dA, dR, 64, 40, 4, 0}, {dA, nullptr, 64, 0, 4, 0},
                                   {dA, dR, 64, 64, 4, 0}};
    CHECK(rz_dispatch_choose(ctx, nullptr, lists, 4, s) == RZ_OK, "rz_dispatch_choose");
    CHECK(rz_dispatch_collect(ctx, choice, 3) == RZ_EINVAL, "rz_dispatch_collect with another list count");
    CHECK(rz_dispatch_collect(ctx, choice, 4) == RZ_OK, "rz_dispatch_collect");
    CHECK(choice[0] == RZ_METHOD_REDZONE && choice[1] == RZ_METHOD_PARTITION && choice[2] == RZ_METHOD_REDZONE &&
              choice[3] == RZ_METHOD_REDZONE,
          "the rule's choices (k = 0 and k = n: Redzone)");
    CHECK(rz_dispatch_collect(ctx, choice, 4) == RZ_EINVAL, "a choice is collected once");
    CHECK(rz_remove_auto_batched(ctx, nullptr, nullptr, 2, 1 << 22, 0, nullptr, s) == RZ_EINVAL,
          "rz_remove_auto_batched without lists");
    CHECK(rz_remove_auto_batched(ctx, nullptr, lists, 4, -1, 0, nullptr, s) == RZ_EINVAL,
          "rz_remove_auto_batched with a negative threshold");
    CHECK(rz_remove_auto_batched(ctx, nullptr, bad, 1, 0, 0, nullptr, s) == RZ_EINVAL,
          "rz_remove_auto_batched with k > n (a large list)");
    HIPCHECK(hipFree(dA));
    HIPCHECK(hipFree(dR));
  }

  HIPCHECK(hipStreamDestroy(s));
  CHECK(rz_ctx_destroy(ctx) == RZ_OK, "rz_ctx_destroy");
  if (failures) {
    printf("FAILED: %d check(s)\n", failures);
    return 1;
  }
  printf("PASS\n");
  return 0;
}
