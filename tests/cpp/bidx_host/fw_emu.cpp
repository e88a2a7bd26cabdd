// Floyd-Warshall for the host check's stand-in of the batched closure,
// built without sanitizers (it only closes the index's part matrices).
#include <stdint.h>
#include <algorithm>
struct fap_ctx;
namespace fapi { constexpr int32_t kInf = 0x3FFFFFFF; }
enum { FAP_OK = 0 };
extern "C" int fap_fw_batched(fap_ctx *, int32_t count, int32_t *const *m, const int32_t *ns, const int64_t *lds, void *) {
  for (int q = 0; q < count; q++) {
    const int n = ns[q]; const int64_t ld = lds[q]; int32_t *a = m[q];
    for (int k = 0; k < n; k++) for (int i = 0; i < n; i++) {
      const int aik = a[i * ld + k];
      if (aik >= fapi::kInf) continue;
      for (int j = 0; j < n; j++) a[i * ld + j] = std::min(a[i * ld + j], aik + a[k * ld + j]);
    }
  }
  return FAP_OK;
}

