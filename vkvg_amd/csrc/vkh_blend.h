// UNORM8 premultiplied compositing of one pixel, shared by the HIP kernels and the CPU test that
// checks it exhaustively (tests/test_blend.py). Operators of src/vkvg_device_internal.c:200-218:
// OVER / SOURCE d' = s + d (1 - s.a), DIFFERENCE d' = s - d (1 - s.a) clamped at 0 (blend op
// SUBTRACT), CLEAR (logic op) 0. d (1 - s.a) per channel is the exact rounding of d*(255-sa)/255.
#ifndef VKH_BLEND_H
#define VKH_BLEND_H

#include <stdint.h>

#ifdef __HIPCC__
#define VKH_HD __host__ __device__ __forceinline__
#else
#define VKH_HD static inline
#endif

// rounding of a*b/255 for a, b <= 255
VKH_HD uint32_t mul255(uint32_t a, uint32_t b) {
    uint32_t t = a * b + 128u;
    return (t + (t >> 8)) >> 8;
}

// reference form: one channel at a time
VKH_HD uint32_t blend_px_channels(uint32_t op, uint32_t d, uint32_t s) {
    if (op == 0)
        return 0;
    uint32_t ia = 255u - (s >> 24), r = 0;
    for (int i = 0; i < 4; i++) {
        uint32_t dc = (d >> (8 * i)) & 0xFFu, sc = (s >> (8 * i)) & 0xFFu, m = mul255(dc, ia), v;
        if (op == 3)
            v = sc > m ? sc - m : 0u;
        else
            v = (sc + m) < 255u ? sc + m : 255u;
        r |= v << (8 * i);
    }
    return r;
}

// two channels at once, each in the low byte of a 16-bit lane (no lane ever carries into the
// next: x*ia + 128 <= 65153 and t + (t >> 8) <= 65407)
VKH_HD uint32_t mul255x2(uint32_t x, uint32_t ia) {
    uint32_t t = x * ia + 0x00800080u;
    return ((t + ((t >> 8) & 0x00FF00FFu)) >> 8) & 0x00FF00FFu;
}

// same result as blend_px_channels for every input
VKH_HD uint32_t blend_px(uint32_t op, uint32_t d, uint32_t s) {
    if (op == 0)
        return 0;
    const uint32_t ia = 255u - (s >> 24);
    const uint32_t m_lo = mul255x2(d & 0x00FF00FFu, ia), m_hi = mul255x2((d >> 8) & 0x00FF00FFu, ia);
    const uint32_t s_lo = s & 0x00FF00FFu, s_hi = (s >> 8) & 0x00FF00FFu;
    uint32_t v_lo, v_hi;
    if (op == 3) { // 256 + s - m keeps bit 8 exactly when s >= m
        const uint32_t x_lo = (s_lo | 0x01000100u) - m_lo, x_hi = (s_hi | 0x01000100u) - m_hi;
        v_lo = x_lo & (((x_lo >> 8) & 0x00010001u) * 0xFFu);
        v_hi = x_hi & (((x_hi >> 8) & 0x00010001u) * 0xFFu);
    } else { // s + m <= 510: bit 8 marks saturation
        const uint32_t x_lo = s_lo + m_lo, x_hi = s_hi + m_hi;
        v_lo = (x_lo | (((x_lo >> 8) & 0x00010001u) * 0xFFu)) & 0x00FF00FFu;
        v_hi = (x_hi | (((x_hi >> 8) & 0x00010001u) * 0xFFu)) & 0x00FF00FFu;
    }
    return v_lo | (v_hi << 8);
}

// The same blend on 16-bit lanes (GCC vector extension, understood by gcc and clang): lanes hold
// channels (0, 2) and (1, 3); on gfx9 each operation is one packed instruction (v_pk_mad_u16,
// v_pk_lshrrev_b16, v_pk_add_u16 with clamp). One sequence serves every operator:
//   m = round(d * ia / 255) is the high byte of u = t + (t >> 8), t = d * ia + 128 (u <= 65407);
//   with a source byte b in the HIGH byte of a lane (b * 256), the saturating 16-bit add
//   b*256 + u = (b + m)*256 + (u & 255) overflows exactly when b + m > 255, so its high byte is
//   min(b + m, 255);
//   OVER / SOURCE: b = s, d' = min(s + m, 255);
//   DIFFERENCE: b = 255 - s and the result inverted: 255 - min(255 - s + m, 255) = max(s - m, 0);
//   CLEAR: b = 255 and the result inverted: 255 - min(255 + m, 255) = 0.
// The result gathers the four high bytes and applies the inversion mask x.
typedef uint16_t vkh_u16x2 __attribute__((vector_size(4)));

VKH_HD vkh_u16x2 vkh_u16x2_of(uint32_t v) {
    vkh_u16x2 r;
    __builtin_memcpy(&r, &v, 4);
    return r;
}
VKH_HD uint32_t vkh_u32_of(vkh_u16x2 v) {
    uint32_t r;
    __builtin_memcpy(&r, &v, 4);
    return r;
}
#ifdef __clang__
#define VKH_ADDSAT_U16X2(a, b) __builtin_elementwise_add_sat(a, b)
#else
#define VKH_ADDSAT_U16X2(a, b) ((a) + (b) | (vkh_u16x2)((vkh_u16x2)((a) + (b)) < (a)))
#endif

// source-dependent operands of blend_px16, computed once for a uniform source
typedef struct {
    vkh_u16x2 ia, b_lo, b_hi; // 255 - alpha in both lanes; bytes b of channels (0, 2) and (1, 3) times 256
    uint32_t x;               // inversion mask of the result
} vkh_src16;

VKH_HD vkh_src16 vkh_src16_of(uint32_t op, uint32_t s) {
    vkh_src16 r;
    const int add = op == 1 || op == 2;
    const uint32_t b = add ? s : op == 3 ? ~s : 0xFFFFFFFFu;
    r.x = add ? 0u : 0xFFFFFFFFu;
#if defined(__HIP_DEVICE_COMPILE__)
    // (255 - a) in the low byte of both lanes; channels 0 and 2 moved into the high bytes
    r.ia = vkh_u16x2_of(__builtin_amdgcn_perm(0u, ~s, 0x0c030c03u));
    r.b_lo = vkh_u16x2_of(__builtin_amdgcn_perm(0u, b, 0x020c000cu));
#else
    const uint32_t ia = 255u - (s >> 24);
    r.ia = vkh_u16x2_of(ia | (ia << 16));
    r.b_lo = vkh_u16x2_of((b & 0x00FF00FFu) << 8);
#endif
    r.b_hi = vkh_u16x2_of(b & 0xFF00FF00u);
    return r;
}

// same result as blend_px_channels(op, d, s) for every input, S = vkh_src16_of(op, s)
VKH_HD uint32_t blend_px16(uint32_t d, vkh_src16 S) {
    const vkh_u16x2 h = {128, 128};
    const vkh_u16x2 t_lo = vkh_u16x2_of(d & 0x00FF00FFu) * S.ia + h;
#if defined(__HIP_DEVICE_COMPILE__)
    // bytes 1 and 3 of d into the low bytes of the 16-bit lanes in one v_perm_b32
    const vkh_u16x2 t_hi = vkh_u16x2_of(__builtin_amdgcn_perm(0u, d, 0x0c030c01u)) * S.ia + h;
#else
    const vkh_u16x2 t_hi = vkh_u16x2_of((d >> 8) & 0x00FF00FFu) * S.ia + h;
#endif
    const vkh_u16x2 v_lo = VKH_ADDSAT_U16X2(S.b_lo, t_lo + (t_lo >> 8));
    const vkh_u16x2 v_hi = VKH_ADDSAT_U16X2(S.b_hi, t_hi + (t_hi >> 8));
#if defined(__HIP_DEVICE_COMPILE__)
    // high bytes: channel 0 = byte 1 of v_lo, 1 = byte 1 of v_hi, 2 = byte 3 of v_lo, 3 = byte 3 of v_hi
    return __builtin_amdgcn_perm(vkh_u32_of(v_hi), vkh_u32_of(v_lo), 0x07030501u) ^ S.x;
#else
    return (((vkh_u32_of(v_lo) >> 8) & 0x00FF00FFu) | (vkh_u32_of(v_hi) & 0xFF00FF00u)) ^ S.x;
#endif
}

#endif
