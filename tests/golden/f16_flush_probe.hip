// Device probe that wrote tests/golden/f16_flush_mi355x.npz: v_cvt_f16_f32 on MI355X with
// half denormals flushed (MODE.FP_DENORM[f16/f64] = 0), on float values bracketing the
// smallest normal half 2^-14 (ties included).  Build: hipcc --offload-arch=gfx950 -O2.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <cstdint>
#include <vector>
#include <cmath>
__global__ void K(const float* in, uint16_t* out, int n) {
    __builtin_amdgcn_s_setreg(2433, 0);  // hwreg(HW_REG_MODE, 6, 2) = 0: flush f16/f64 denormals
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        _Float16 h = static_cast<_Float16>(in[i]);
        out[i] = __builtin_bit_cast(uint16_t, h);
    }
}
int main() {
    std::vector<float> v;
    const float t = std::ldexp(1.0f, -14);
    for (int k = -40; k <= 40; ++k) { float x = std::nextafter(t, k < 0 ? 0.0f : 1.0f); float y = t; for (int j = 0; j < std::abs(k); ++j) y = std::nextafter(y, k < 0 ? 0.0f : 1.0f); v.push_back(y); v.push_back(-y); }
    v.push_back(std::ldexp(1.0f, -20)); v.push_back(std::ldexp(1.0f, -24)); v.push_back(t * (1.0f - 1.0f / 4096)); v.push_back(t * (1.0f - 1.0f / 2048));
    v.push_back(t * (1.0f - 1.0f / 2048) * (1.0f + 1e-7f));
    int n = v.size();
    float* din; uint16_t* dout;
    hipMalloc(&din, n * 4); hipMalloc(&dout, n * 2);
    hipMemcpy(din, v.data(), n * 4, hipMemcpyHostToDevice);
    K<<<(n + 63) / 64, 64>>>(din, dout, n);
    std::vector<uint16_t> o(n);
    hipMemcpy(o.data(), dout, n * 2, hipMemcpyDeviceToHost);
    int flushed_above = 0, kept_below = 0;
    for (int i = 0; i < n; ++i) {
        _Float16 r = static_cast<_Float16>(v[i]);  // host RNE, subnormals kept
        uint16_t rb; std::memcpy(&rb, &r, 2);
        bool rounded_sub = (rb & 0x7C00) == 0;
        uint16_t after = rounded_sub ? (rb & 0x8000) : rb;          // flush if the rounded result is subnormal
        bool exact_sub = std::fabs(v[i]) < t;
        uint16_t before = exact_sub ? (rb & 0x8000) : rb;            // flush if the exact value is below 2^-14
        printf("%+.9e dev %04x after %04x before %04x %s\n", v[i], o[i], after, before,
               o[i] == after && o[i] == before ? "both" : (o[i] == after ? "AFTER" : (o[i] == before ? "BEFORE" : "NEITHER")));
    }
    hipFree(din); hipFree(dout);
    return 0;
}
