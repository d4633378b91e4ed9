// layout.hip -- host-side conversion between this library's key and BRT
// records and the reference's byte layouts:
//   BigUnsigned (C/BigUnsigned.h:8-16): int len, then 11 8-bit blocks,
//     least significant first, with no zero block at len - 1 (the form
//     zapLeadingZeros leaves, C/BigUnsigned.c:22-26); 16 bytes.  Blocks at
//     and above len are not part of the value (getBUBlock reads them as 0).
//   BrtNode (C/BrtNode.h:10-21): left, left_leaf, right_leaf, lcp
//     (BigUnsigned), lcp_length, parent; 32 bytes.
// A caller that reads device results back into the reference's structs (as
// tests/KernelTests.cpp:74,130,190 do) converts with these.
#include <cstring>
#include <string>

#include "../../include/pgvd.h"
#include "capi_internal.hpp"

using namespace pgvd_capi;

namespace {

static_assert(sizeof(pgvd_big_unsigned) == 16, "BigUnsigned is 16 bytes");
static_assert(sizeof(pgvd_ref_brt_node) == 32, "BrtNode is 32 bytes");
static_assert(offsetof(pgvd_ref_brt_node, lcp) == 8 && offsetof(pgvd_ref_brt_node, lcp_length) == 24 &&
                  offsetof(pgvd_ref_brt_node, parent) == 28,
              "BrtNode field offsets");

constexpr int kBlocks = 11;   // BIG_INTEGER_SIZE (C/BigUnsigned.h:8)

// (lo, hi) -> normalised BigUnsigned; false when the value needs more than
// 88 bits.  Unused blocks are zeroed.
bool to_big(uint64_t lo, uint64_t hi, pgvd_big_unsigned *b) {
  if (hi >> 24) return false;
  std::memset(b, 0, sizeof(*b));
  int len = 0;
  for (int i = 0; i < kBlocks; ++i) {
    const uint8_t byte = (uint8_t)(i < 8 ? lo >> (8 * i) : hi >> (8 * (i - 8)));
    b->blk[i] = byte;
    if (byte) len = i + 1;
  }
  b->len = len;
  return true;
}

// BigUnsigned -> (lo, hi); false when len is outside [0, 11].
bool from_big(const pgvd_big_unsigned *b, uint64_t *lo, uint64_t *hi) {
  if (b->len < 0 || b->len > kBlocks) return false;
  uint64_t l = 0, h = 0;
  for (int i = 0; i < b->len; ++i) {
    if (i < 8) l |= (uint64_t)b->blk[i] << (8 * i);
    else h |= (uint64_t)b->blk[i] << (8 * (i - 8));
  }
  *lo = l;
  *hi = h;
  return true;
}

void read_key(const void *keys, int kb, int64_t i, uint64_t *lo, uint64_t *hi) {
  if (kb == 4) {
    *lo = static_cast<const uint32_t *>(keys)[i];
    *hi = 0;
  } else if (kb == 8) {
    *lo = static_cast<const uint64_t *>(keys)[i];
    *hi = 0;
  } else {
    *lo = static_cast<const uint64_t *>(keys)[2 * i];
    *hi = static_cast<const uint64_t *>(keys)[2 * i + 1];
  }
}

bool write_key(void *keys, int kb, int64_t i, uint64_t lo, uint64_t hi) {
  if (kb == 4) {
    if (hi || (lo >> 32)) return false;
    static_cast<uint32_t *>(keys)[i] = (uint32_t)lo;
  } else if (kb == 8) {
    if (hi) return false;
    static_cast<uint64_t *>(keys)[i] = lo;
  } else {
    static_cast<uint64_t *>(keys)[2 * i] = lo;
    static_cast<uint64_t *>(keys)[2 * i + 1] = hi;
  }
  return true;
}

// The library's BRT record of key width kb (pgvd_brt_node32/64/128).
size_t brt_record_bytes(int kb) {
  return kb == 4 ? sizeof(pgvd_brt_node32) : kb == 8 ? sizeof(pgvd_brt_node64) : sizeof(pgvd_brt_node128);
}

int check_kb(int kb) {
  if (kb != 4 && kb != 8 && kb != 16) return fail(PGVD_ERR_INVALID, "key_bytes must be 4, 8 or 16");
  return PGVD_OK;
}

}  // namespace

extern "C" {

int pgvd_keys_to_big_unsigned(const void *keys, int key_bytes, int64_t n, pgvd_big_unsigned *out) {
  CHK(check_kb(key_bytes));
  if (n < 0 || (n > 0 && (!keys || !out))) return fail(PGVD_ERR_INVALID, "null argument");
  for (int64_t i = 0; i < n; ++i) {
    uint64_t lo, hi;
    read_key(keys, key_bytes, i, &lo, &hi);
    if (!to_big(lo, hi, &out[i]))
      return fail(PGVD_ERR_TOO_LARGE, "key " + std::to_string(i) + " needs more than 88 bits");
  }
  return PGVD_OK;
}

int pgvd_big_unsigned_to_keys(const pgvd_big_unsigned *in, int64_t n, int key_bytes, void *keys) {
  CHK(check_kb(key_bytes));
  if (n < 0 || (n > 0 && (!keys || !in))) return fail(PGVD_ERR_INVALID, "null argument");
  for (int64_t i = 0; i < n; ++i) {
    uint64_t lo, hi;
    if (!from_big(&in[i], &lo, &hi))
      return fail(PGVD_ERR_INVALID, "BigUnsigned " + std::to_string(i) + " has len outside 0..11");
    if (!write_key(keys, key_bytes, i, lo, hi))
      return fail(PGVD_ERR_TOO_LARGE, "BigUnsigned " + std::to_string(i) + " does not fit " +
                                          std::to_string(key_bytes) + "-byte keys");
  }
  return PGVD_OK;
}

int pgvd_compare_big_unsigned(const pgvd_big_unsigned *x, const pgvd_big_unsigned *y) {
  // compareBU (C/BigUnsigned.c:132-153): the longer is larger; equal
  // lengths compare block by block from the most significant one
  if (x->len != y->len) return x->len < y->len ? -1 : 1;
  for (int i = x->len - 1; i >= 0; --i)
    if (x->blk[i] != y->blk[i]) return x->blk[i] < y->blk[i] ? -1 : 1;
  return 0;
}

int pgvd_brt_to_ref_nodes(const void *brt, const int32_t *parent, int64_t count, int key_bytes,
                          pgvd_ref_brt_node *out) {
  CHK(check_kb(key_bytes));
  if (count < 0 || (count > 0 && (!brt || !parent || !out))) return fail(PGVD_ERR_INVALID, "null argument");
  const char *rec = static_cast<const char *>(brt);
  const size_t rb = brt_record_bytes(key_bytes);
  for (int64_t i = 0; i < count; ++i, rec += rb) {
    uint64_t lo, hi = 0;
    int32_t left;
    uint16_t len;
    uint8_t ll, rl;
    if (key_bytes == 4) {
      const pgvd_brt_node32 *r = reinterpret_cast<const pgvd_brt_node32 *>(rec);
      lo = r->lcp; left = r->left; len = r->lcp_length; ll = r->left_leaf; rl = r->right_leaf;
    } else if (key_bytes == 8) {
      const pgvd_brt_node64 *r = reinterpret_cast<const pgvd_brt_node64 *>(rec);
      lo = r->lcp; left = r->left; len = r->lcp_length; ll = r->left_leaf; rl = r->right_leaf;
    } else {
      const pgvd_brt_node128 *r = reinterpret_cast<const pgvd_brt_node128 *>(rec);
      lo = r->lcp_lo; hi = r->lcp_hi; left = r->left; len = r->lcp_length; ll = r->left_leaf;
      rl = r->right_leaf;
    }
    pgvd_ref_brt_node &o = out[i];
    std::memset(&o, 0, sizeof(o));
    o.left = left;
    o.left_leaf = ll ? 1 : 0;
    o.right_leaf = rl ? 1 : 0;
    if (!to_big(lo, hi, &o.lcp))
      return fail(PGVD_ERR_TOO_LARGE, "BRT node " + std::to_string(i) + ": lcp needs more than 88 bits");
    o.lcp_length = len;
    o.parent = parent[i];
  }
  return PGVD_OK;
}

int pgvd_ref_nodes_to_brt(const pgvd_ref_brt_node *in, int64_t count, int key_bytes, void *brt,
                          int32_t *parent) {
  CHK(check_kb(key_bytes));
  if (count < 0 || (count > 0 && (!brt || !parent || !in))) return fail(PGVD_ERR_INVALID, "null argument");
  char *rec = static_cast<char *>(brt);
  const size_t rb = brt_record_bytes(key_bytes);
  for (int64_t i = 0; i < count; ++i, rec += rb) {
    const pgvd_ref_brt_node &r = in[i];
    uint64_t lo, hi;
    if (!from_big(&r.lcp, &lo, &hi))
      return fail(PGVD_ERR_INVALID, "BrtNode " + std::to_string(i) + ": lcp len outside 0..11");
    if (r.lcp_length < 0 || r.lcp_length > 128)
      return fail(PGVD_ERR_INVALID, "BrtNode " + std::to_string(i) + ": lcp_length outside 0..128");
    if ((key_bytes == 4 && (hi || lo >> 32)) || (key_bytes == 8 && hi))
      return fail(PGVD_ERR_TOO_LARGE, "BrtNode " + std::to_string(i) + ": lcp does not fit the key width");
    if (key_bytes == 4) {
      pgvd_brt_node32 *o = reinterpret_cast<pgvd_brt_node32 *>(rec);
      std::memset(o, 0, sizeof(*o));
      o->lcp = (uint32_t)lo; o->left = r.left; o->lcp_length = (uint16_t)r.lcp_length;
      o->left_leaf = r.left_leaf ? 1 : 0; o->right_leaf = r.right_leaf ? 1 : 0;
    } else if (key_bytes == 8) {
      pgvd_brt_node64 *o = reinterpret_cast<pgvd_brt_node64 *>(rec);
      std::memset(o, 0, sizeof(*o));
      o->lcp = lo; o->left = r.left; o->lcp_length = (uint16_t)r.lcp_length;
      o->left_leaf = r.left_leaf ? 1 : 0; o->right_leaf = r.right_leaf ? 1 : 0;
    } else {
      pgvd_brt_node128 *o = reinterpret_cast<pgvd_brt_node128 *>(rec);
      std::memset(o, 0, sizeof(*o));
      o->lcp_lo = lo; o->lcp_hi = hi; o->left = r.left; o->lcp_length = (uint16_t)r.lcp_length;
      o->left_leaf = r.left_leaf ? 1 : 0; o->right_leaf = r.right_leaf ? 1 : 0;
    }
    parent[i] = r.parent;
  }
  return PGVD_OK;
}

}  // extern "C"
