// Device layout of an uploaded scene (shared by the upload code and the
// kernels).
//
// Inner node record (one per FastNode with children), RECW(M) 32-bit words,
// 128/256/384/512 bytes for M = 4/8/12/16, 64-byte aligned:
//   [0, 6M)        child boxes, SoA planes of M floats:
//                  minX minY minZ maxX maxY maxZ (slot s of plane p at p*M+s).
//                  Slots [childCount, k) are the reference's zero boxes,
//                  slots >= k are never read.
//   [6M, ...)      header block, the fields
//                  innerBase  record index of the first inner child
//                  leafBase   leaf ordinal of the first leaf child
//                  leafMask   bit s: child slot s is a leaf
//                  c          childCount
//                  kcodes     2 bits per child slot: (k_child / 4) - 1
//                  childBase  FastNode id of child slot 0 (childIdBegin)
//                  rows       traverseOrderEachAxis, 4-bit slots per axis
//                  packed so that a lane reads what its visit needs with
//                  one 16-byte load when M <= 8 (indices below 2^28):
//    M = 4:  w0 innerBase | leafMask << 28
//            w1 leafBase | (c - 1) << 28
//            w2 kcodes | row_x << 8 | (row_y & 0xFF) << 24
//            w3 row_y >> 8 | row_z << 8   (row_a = bits 8+16a.. of w3:w2)
//            w4 childBase
//    M = 8:  one block per axis a, w[4a .. 4a+3]:
//            innerBase, leafBase, leafMask | c << 8 | kcodes << 16, row_a
//            w12 childBase
//    M > 8:  w0 innerBase, w1 leafBase, w2 leafMask | c << 16,
//            w3 kcodes, w4 childBase, w5.. rows (OW = ceil(M/8) words
//            per axis)
// Children of one FastNode are contiguous (customTreeOrder), so the inner
// children get consecutive record indices and the leaf children consecutive
// leaf ordinals: the stack entry of child slot s is computed from
// innerBase/leafBase and a popcount of leafMask, without a table lookup.
//
// Quantized inner node record (the per-ray closest/any kernels; built when
// the scene qualifies, see below), QRECW(M) words, 128 bytes for M = 4, 8
// and 192 for M = 12, 16:
//   w0-5   the node's own box as its parent tests it (minX minY minZ maxX
//          maxY maxZ, exact floats; the root: the union of its children)
//   w6     byte a (a = 0, 1, 2): e_a + 160, the frame scale s_a = 2^e_a
//   w7     childBase
//   w8..   header: M = 4 the words w0-w3 of the exact record's header;
//          M = 8 the three per-axis blocks; M > 8 innerBase, leafBase,
//          leafMask | c << 16, kcodes, then the rows (2 words per axis)
//   planes per axis a: the lo bytes of the M child slots, then the hi
//          bytes (M = 4: 4 + 4 bytes per axis, the three axes in 24 bytes;
//          M = 8: 8 + 8; M > 8: 16 + 16, zero padded).  A real child's
//          exact plane lies in the quantized one's frame:
//            F_a + q_lo * s_a <= lo   and   F_a + q_hi * s_a >= hi
//          (real arithmetic, F = the own box's min), so a box test on the
//          quantized planes, with a rounding margin, passes whenever the
//          reference's test of the exact box passes.
// A scene qualifies when every box coordinate is within +-2^60, every real
// child box has lo <= hi and lies inside its parent's own box, and every
// padded slot [c, k) is the zero box.
//
// Leaf block (one per leaf, in FastNode order == trianglePoints order):
//   LSTRIDE(Lm) = 8 + 20 * ceil(Lm / 2) floats.  The first 8 floats are the
//   leaf's head: its own box as the parent tests it (minX minY minZ maxX
//   maxY maxZ), then 2 floats of padding.  Then the leaf's slots in pairs,
//   one 80-byte chunk per pair (slots 2c and 2c+1).  A chunk holds, for each of
//   the nine values p0.x p0.y p0.z e1.x e1.y e1.z e2.x e2.y e2.z, the two
//   slots' floats side by side (18 floats), then 2 floats of padding, so a
//   pair of triangles is four 16-byte loads and one 8-byte load.  e1 = p1 -
//   p0 and e2 = p2 - p0 are the FP32 differences the intersection kernels
//   form first.  Slots >= k (the leaf's loop count) are zero.  Word 18 of
//   chunk 0 (a uint32) is the number of leading slots to test for a ray
//   with a finite direction: up to the last triangle that is not all
//   zeros, rounded up to a pair (an all-zero triangle has divisor 0 for
//   every finite direction and never passes the test).
//
// Tentative stack entries (quantized records): a child whose quantized box
// passed is pushed with bit 27 (kEntryTent) set, and every entry is pushed
// with the ray's tmax at the push beside it; the pop of a tentative entry
// tests its exact box (the record's or leaf head's own box) with that tmax
// and skips the visit when the exact test fails.  Entries pushed for certain
// (stale aabbDistances slots, padded slots, the root) have the bit clear.
//
// Stack entry (32 bits):
//   bit 31        leaf
//   bit 30        invalid (FastNode id outside the node array)
//   bits 28-29    kcode: loop count k = (kcode + 1) * 4 of that node/leaf
//   bits 0-27     inner record index / leaf ordinal
#pragma once
#include <cstdint>

namespace rt {

constexpr uint32_t kEntryLeaf = 0x80000000u;
constexpr uint32_t kEntryInvalid = 0x40000000u;
constexpr uint32_t kEntryIndexMask = 0x0FFFFFFFu;
// quantized records: bit 27 marks a tentative entry, the index has 27 bits
constexpr uint32_t kEntryTent = 0x08000000u;
constexpr uint32_t kEntryIndexMaskQ = 0x07FFFFFFu;
constexpr int kEntryKShift = 28;

__host__ __device__ constexpr uint32_t order_words(uint32_t M) { return (M + 7) / 8; }
__host__ __device__ constexpr uint32_t header_words(uint32_t M) {
    return M == 4 ? 5 : (M == 8 ? 13 : 5 + 3 * order_words(M));
}
// Header word holding childBase.
__host__ __device__ constexpr uint32_t child_base_word(uint32_t M) { return M == 8 ? 12 : 4; }
__host__ __device__ constexpr uint32_t record_words(uint32_t M) {
    return ((6 * M + header_words(M)) + 15) / 16 * 16;
}
constexpr uint32_t kLeafHead = 8;  // floats before a leaf block's slot pairs
__host__ __device__ constexpr uint32_t leaf_stride(uint32_t Lm) { return kLeafHead + (Lm + 1) / 2 * 20; }
// Float offset of value v (0..8: p0 xyz, e1 xyz, e2 xyz) of slot s in a leaf block.
__host__ __device__ constexpr uint32_t leaf_offset(uint32_t s, uint32_t v) {
    return kLeafHead + (s >> 1) * 20 + 2 * v + (s & 1);
}

// Quantized record geometry.
__host__ __device__ constexpr uint32_t qrecord_words(uint32_t M) { return M <= 8 ? 32 : 48; }
constexpr uint32_t kQHeader = 8;  // header word
__host__ __device__ constexpr uint32_t qplanes_word(uint32_t M) { return kQHeader + (M == 4 ? 4 : 12); }
// byte offset of the lo (hi = 1) bytes of axis a in the planes block
__host__ __device__ constexpr uint32_t qplane_byte(uint32_t M, uint32_t a, uint32_t hi) {
    return M == 4 ? 8 * a + 4 * hi : (M == 8 ? 16 * a + 8 * hi : 32 * a + 16 * hi);
}
constexpr int kQExpBias = 160;
constexpr float kQMaxCoord = 0x1p60f;  // box and ray bound of the quantized test

}  // namespace rt
