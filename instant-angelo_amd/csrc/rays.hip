// Training-ray sampling on the device (systems/neus.py:35-118, train stage):
// per ray k a view v = index[k] and a pixel (y[k], x[k]) drawn by the caller;
// the camera-space direction of that pixel rotated by the view's c2w
// (models/ray_utils.py:23-31: rays_d_i = sum_j d_j c2w[i][j]), normalized as
// F.normalize does (x / max(|x|, 1e-12)); the origin is the c2w translation;
// the pixel's colour and foreground mask gathered, and, when the dataset
// applies masks, blended with the background colour.  Replaces the host-side
// index / gather / get_rays round trip of the reference's preprocess_data.
//
// One lane per ray: the work is a handful of dependent gathers (c2w row
// triplet, a direction, a pixel) and 40 B written per ray.
#include <cmath>
#include <rocprim/device/device_radix_sort.hpp>
#include "am_common.h"

namespace am {

struct RaySampleArgs {
    const float* images;      // [n_views, H, W, C], C >= 3 (channels 0..2 used)
    const float* fg_masks;    // [n_views, H, W] or nullptr (all foreground)
    const float* c2w;         // [n_views, 3, 4]
    const float* directions;  // [H, W, 3] or [n_views, H, W, 3]
    const float* bg_color;    // [3] device, or nullptr when apply_mask == 0
    int32_t n_views, H, W, C, dirs_per_view, apply_mask;
};

__global__ void k_sample_rays(RaySampleArgs a, const int32_t* __restrict__ index,
                              const int32_t* __restrict__ ys, const int32_t* __restrict__ xs,
                              int64_t n, float* __restrict__ rays, float* __restrict__ rgb,
                              float* __restrict__ fg, int32_t* __restrict__ bad) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int v = index[k], y = ys[k], x = xs[k];
    // out-of-range draws are clamped (no out-of-bounds read) and counted
    if (v < 0 || v >= a.n_views || y < 0 || y >= a.H || x < 0 || x >= a.W) {
        atomicAdd(bad, 1);
        v = min(max(v, 0), a.n_views - 1);
        y = min(max(y, 0), a.H - 1);
        x = min(max(x, 0), a.W - 1);
    }
    const int64_t pix = ((int64_t)v * a.H + y) * a.W + x;
    const int64_t dpix = a.dirs_per_view ? pix : (int64_t)y * a.W + x;
    const float d0 = a.directions[3 * dpix], d1 = a.directions[3 * dpix + 1],
                d2 = a.directions[3 * dpix + 2];
    const float* m = a.c2w + 12 * (int64_t)v;
    float r[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) r[i] = (d0 * m[4 * i] + d1 * m[4 * i + 1]) + d2 * m[4 * i + 2];
    const float nrm = sqrtf((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]);
    const float den = fmaxf(nrm, 1e-12f);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        rays[6 * k + i] = m[4 * i + 3];
        rays[6 * k + 3 + i] = r[i] / den;
    }
    const float f = a.fg_masks ? a.fg_masks[pix] : 1.0f;
    const float* px = a.images + pix * a.C;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float c = px[i];
        if (a.apply_mask) c = c * f + a.bg_color[i] * (1.0f - f);
        rgb[3 * k + i] = c;
    }
    fg[k] = f;
}

// Morton (Z-order) key of a sample: its position in [-r, r]^3 quantised to
// `bits` (<= 10) bits per axis and the bits interleaved (x highest).  Sorting samples by
// it makes the waves that run together gather nearby table rows.
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void k_morton_keys(const float* __restrict__ pts, int64_t n, float r, float inv2r,
                              int bits, int32_t* __restrict__ keys) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t cells = 1u << bits;
    uint32_t q[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        float u = (pts[3 * i + d] + r) * inv2r;
        u = fminf(fmaxf(u, 0.0f), 1.0f);   // NaN -> 0
        const uint32_t v = (uint32_t)(u * (float)cells);
        q[d] = v > cells - 1 ? cells - 1 : v;
    }
    keys[i] = (int32_t)((spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]));
}

__global__ void k_iota(int64_t n, int32_t* __restrict__ v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (int32_t)i;
}

__global__ void k_invert(const int32_t* __restrict__ perm, int64_t n, int32_t* __restrict__ inv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) inv[perm[i]] = (int32_t)i;
}

static size_t order_sort_bytes(int64_t n, int bits) {
    size_t bytes = 0;
    rocprim::radix_sort_pairs(nullptr, bytes, (const int32_t*)nullptr, (int32_t*)nullptr,
                              (const int32_t*)nullptr, (int32_t*)nullptr, (size_t)n, 0u,
                              (unsigned)(3 * bits), nullptr, false);
    return bytes;
}

}  // namespace am

using namespace am;

extern "C" size_t am_morton_order_workspace_size(int64_t n, int32_t bits) {
    if (n < 0 || bits < 1 || bits > 10) return 0;
    WsCarver c(nullptr, 0);
    for (int b = 0; b < 3; ++b) c.take<int32_t>((size_t)n + 1);
    c.take<char>(order_sort_bytes(n > 0 ? n : 1, bits) + 256);
    return c.used + 256;
}

extern "C" int am_morton_order(const float* pts, int64_t n, float radius, int32_t bits,
                               int32_t* perm, int32_t* inv, void* workspace,
                               size_t workspace_bytes, void* stream) {
    AM_CHECK_ARG(n >= 0 && n < (1ll << 31), "n %lld", (long long)n);
    AM_CHECK_ARG(radius > 0.0f, "radius %g <= 0", (double)radius);
    AM_CHECK_ARG(bits >= 1 && bits <= 10, "bits %d out of [1, 10]", bits);
    if (n == 0) return AM_OK;
    AM_CHECK_ARG(pts && perm && inv && workspace, "NULL buffer");
    hipStream_t s = to_stream(stream);
    WsCarver c(workspace, workspace_bytes);
    int32_t* keys = c.take<int32_t>((size_t)n + 1);
    int32_t* keys2 = c.take<int32_t>((size_t)n + 1);
    int32_t* vals = c.take<int32_t>((size_t)n + 1);
    size_t sb = order_sort_bytes(n, bits);
    char* tmp = c.take<char>(sb + 256);
    if (!c.ok()) return workspace_short(c.used, workspace_bytes);
    const unsigned nb = (unsigned)ceil_div(n, 256);
    hipLaunchKernelGGL(k_morton_keys, dim3(nb), dim3(256), 0, s, pts, n, radius, 0.5f / radius,
                       (int)bits, keys);
    hipLaunchKernelGGL(k_iota, dim3(nb), dim3(256), 0, s, n, vals);
    AM_LAUNCH_CHECK();
    // a radix sort is stable: equal keys keep the caller's order
    hipError_t e = rocprim::radix_sort_pairs(tmp, sb, (const int32_t*)keys, keys2,
                                             (const int32_t*)vals, perm, (size_t)n, 0u,
                                             (unsigned)(3 * bits), s, false);
    if (e != hipSuccess) {
        set_error("radix sort failed: %s", hipGetErrorString(e));
        return AM_ERR_HIP;
    }
    hipLaunchKernelGGL(k_invert, dim3(nb), dim3(256), 0, s, perm, n, inv);
    AM_LAUNCH_CHECK();
    return AM_OK;
}

extern "C" int am_morton_keys(const float* pts, int64_t n, float radius, int32_t* keys,
                              void* stream) {
    AM_CHECK_ARG(n >= 0, "n < 0");
    AM_CHECK_ARG(radius > 0.0f, "radius %g <= 0", (double)radius);
    if (n == 0) return AM_OK;
    AM_CHECK_ARG(pts && keys, "NULL buffer");
    hipLaunchKernelGGL(k_morton_keys, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                       to_stream(stream), pts, n, radius, 0.5f / radius, 10, keys);
    AM_LAUNCH_CHECK();
    return AM_OK;
}

extern "C" int am_sample_rays(const float* images, const float* fg_masks, const float* c2w,
                              const float* directions, int32_t n_views, int32_t H, int32_t W,
                              int32_t C, int32_t dirs_per_view, const float* bg_color,
                              int32_t apply_mask, const int32_t* index, const int32_t* ys,
                              const int32_t* xs, int64_t n, float* rays, float* rgb,
                              float* fg_mask, int32_t* n_out_of_range, void* stream) {
    AM_CHECK_ARG(n >= 0, "n < 0");
    AM_CHECK_ARG(n_views >= 1 && H >= 1 && W >= 1, "empty image set (%d views, %dx%d)", n_views,
                 W, H);
    AM_CHECK_ARG(C >= 3, "images need at least 3 channels (got %d)", C);
    AM_CHECK_ARG(!apply_mask || bg_color, "apply_mask needs bg_color");
    AM_CHECK_ARG((int64_t)n_views * H * W < (1ll << 40), "image set too large");
    if (n == 0) return AM_OK;
    AM_CHECK_ARG(images && c2w && directions && index && ys && xs && rays && rgb && fg_mask &&
                     n_out_of_range, "NULL buffer");
    RaySampleArgs a;
    a.images = images;
    a.fg_masks = fg_masks;
    a.c2w = c2w;
    a.directions = directions;
    a.bg_color = bg_color;
    a.n_views = n_views;
    a.H = H;
    a.W = W;
    a.C = C;
    a.dirs_per_view = dirs_per_view ? 1 : 0;
    a.apply_mask = apply_mask ? 1 : 0;
    {
        KernelTimer kt_("sample_rays", to_stream(stream));
        hipLaunchKernelGGL(k_sample_rays, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                           to_stream(stream), a, index, ys, xs, n, rays, rgb, fg_mask,
                           n_out_of_range);
    }
    AM_LAUNCH_CHECK();
    return AM_OK;
}
