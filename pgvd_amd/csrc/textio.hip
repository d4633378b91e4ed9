// textio.hip -- the octree's text form on the device: operator<< and
// operator>> for OctNode and std::vector<OctNode> (C/OctNode.h:58-98,
// commented out in the reference).  The text is "N " followed, for every
// node, by each child and the leaf mask in decimal, each followed by one
// space.
//
// Both directions are one pass with a decoupled look-back over tiles:
//   format: a workgroup takes 4 x 256 nodes, sums their text lengths, scans
//           them in the workgroup and across tiles, and prints each 256 into
//           LDS and copies the run out at its offset in turn;
//   parse:  a workgroup takes 4 x 8192 bytes, marks where tokens start
//           (a non-space after a space), scans the counts to token numbers,
//           parses every token from LDS and writes the tile's records
//           k - 1 as one run (token 0 is N, read by the host).
#include "kernels.hpp"
#include "tilescan.hpp"

namespace pgvd {

namespace {

// workspace words (uint64): [0] tile counter (low 32 bits), [1] total
// (bytes or tokens, written by the last tile), [2] first malformed token + 1
// (atomic min over ~0: none), then the tile status array from word 4.
constexpr int kWsCounter = 0, kWsTotal = 1, kWsBad = 2, kWsStatus = 4;
constexpr int kFmtBlock = 256;        // nodes per format tile
constexpr int kMaxNodeText = 9 * 12;  // 3D node: 9 values of at most 11 chars + space
constexpr int kParseBytes = 16;       // bytes per thread in a parse tile
#ifndef PGVD_PARSE_BLOCK
#define PGVD_PARSE_BLOCK 512
#endif
constexpr int kParseBlock = PGVD_PARSE_BLOCK;   // threads per parse workgroup
constexpr int kParseTile = kParseBlock * kParseBytes;   // bytes per round
#ifndef PGVD_PARSE_ROUNDS
#define PGVD_PARSE_ROUNDS 4
#endif
constexpr int kParseRounds = PGVD_PARSE_ROUNDS;          // rounds per tile
static_assert(kParseRounds >= 1 && kParseRounds <= 4, "start counts are packed in 16-bit fields");

__device__ inline int dec_len(int32_t v) {
  uint32_t u = v < 0 ? 0u - (uint32_t)v : (uint32_t)v;
  int n = 1;
  if (u >= 10u) n = 2;
  if (u >= 100u) n = 3;
  if (u >= 1000u) n = 4;
  if (u >= 10000u) n = 5;
  if (u >= 100000u) n = 6;
  if (u >= 1000000u) n = 7;
  if (u >= 10000000u) n = 8;
  if (u >= 100000000u) n = 9;
  if (u >= 1000000000u) n = 10;
  return n + (v < 0);
}

// Prints v and one space at s[at..]; returns the position after the space.
__device__ inline int put_dec(uint8_t *s, int at, int32_t v) {
  const int n = dec_len(v);
  uint32_t u = v < 0 ? 0u - (uint32_t)v : (uint32_t)v;
  int e = at + n;
  s[e] = ' ';
  do {
    s[--e] = (uint8_t)('0' + u % 10u);
    u /= 10u;
  } while (u);
  if (v < 0) s[at] = '-';
  return at + n + 1;
}

__global__ void ws_zero_kernel(uint64_t *ws, int64_t words) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < words;
       i += (int64_t)gridDim.x * kBlock)
    ws[i] = i == kWsBad ? ~0ull : 0ull;
}

#ifndef PGVD_FMT_ROUNDS
#define PGVD_FMT_ROUNDS 4
#endif
constexpr int kFmtRounds = PGVD_FMT_ROUNDS;   // 256-node rounds per tile
static_assert(kFmtRounds >= 1 && kFmtRounds <= 4, "round lengths are packed in 16-bit fields");

// A tile is kFmtRounds rounds of 256 nodes with one look-back: the rounds'
// text lengths are scanned together as 16-bit fields of one word (a round
// holds at most 256 * 108 bytes, below 2^16), then each round is printed
// into LDS and copied out in turn.
template <int DIM>
__global__ __launch_bounds__(kFmtBlock) void format_text_kernel(
    const int32_t *__restrict__ octree, int64_t O, int64_t N_hdr, int hdr_len,
    uint8_t *__restrict__ text, int64_t capacity, uint64_t *__restrict__ ws, int ntiles) {
  constexpr int STRIDE = (1 << DIM) + 1;
  __shared__ uint8_t s_txt[kFmtBlock * kMaxNodeText];
  __shared__ uint64_t s_w[kFmtBlock / 64];
  __shared__ int s_tile;
  const int t = threadIdx.x;
  if (t == 0) s_tile = (int)atomicAdd(reinterpret_cast<unsigned int *>(ws + kWsCounter), 1u);
  __syncthreads();
  const int tile = s_tile;
  const int64_t first = (int64_t)tile * kFmtBlock * kFmtRounds;
  uint64_t packed = 0;
  int32_t cur[STRIDE];   // the records of the round being printed (round 0 kept from here)
#pragma unroll
  for (int r = 0; r < kFmtRounds; ++r) {
    const int64_t i = first + r * kFmtBlock + t;
    uint32_t len = 0;
#pragma unroll
    for (int k = 0; k < STRIDE; ++k) {
      const int32_t x = i < O ? octree[i * STRIDE + k] : 0;
      if (r == 0) cur[k] = x;
      len += (uint32_t)dec_len(x) + 1u;
    }
    packed |= (uint64_t)(i < O ? len : 0u) << (16 * r);
  }
  uint64_t tot_packed;
  const uint64_t ex = block_excl_sum<uint64_t, kFmtBlock / 64>(packed, s_w, &tot_packed);
  uint32_t total = 0;
#pragma unroll
  for (int r = 0; r < kFmtRounds; ++r) total += (uint32_t)(tot_packed >> (16 * r)) & 0xffffu;
  const uint64_t before = tile_prefix<kFmtBlock>(ws + kWsStatus, tile, total);
  if (tile == ntiles - 1 && t == 0) ws[kWsTotal] = (uint64_t)hdr_len + before + total;
  if (!text) return;
  int64_t g0 = hdr_len + (int64_t)before;
  if (g0 + (int64_t)total > capacity) return;
  if (tile == 0 && t == 0) {
    // "N " at the start: N printed from the least significant digit
    uint64_t u = (uint64_t)N_hdr;
    int e = hdr_len - 1;
    text[e] = ' ';
    do {
      text[--e] = (uint8_t)('0' + u % 10u);
      u /= 10u;
    } while (u);
  }
#pragma unroll
  for (int r = 0; r < kFmtRounds; ++r) {
    const uint32_t rtotal = (uint32_t)(tot_packed >> (16 * r)) & 0xffffu;
    const int64_t i = first + r * kFmtBlock + t;
    // the next round's records load while this round prints
    int32_t nxt[STRIDE];
    const int64_t in = i + kFmtBlock;
#pragma unroll
    for (int k = 0; k < STRIDE; ++k)
      nxt[k] = (r + 1 < kFmtRounds && in < O) ? octree[in * STRIDE + k] : 0;
    if (r > 0) __syncthreads();   // the previous round's copy has read s_txt
    if (i < O) {
      int at = (int)((ex >> (16 * r)) & 0xffffu);
#pragma unroll
      for (int k = 0; k < STRIDE; ++k) at = put_dec(s_txt, at, cur[k]);
    }
    __syncthreads();
    // the round's run: bytes up to the first 4-byte boundary, whole words, tail
    uint8_t *dst = text + g0;
    int head = (int)((4 - ((uintptr_t)dst & 3)) & 3);
    if (head > (int)rtotal) head = (int)rtotal;
    if (t < head) dst[t] = s_txt[t];
    const int nw = ((int)rtotal - head) >> 2;
    uint32_t *dw = reinterpret_cast<uint32_t *>(dst + head);
    for (int w = t; w < nw; w += kFmtBlock) {
      const int b = head + 4 * w;
      dw[w] = (uint32_t)s_txt[b] | ((uint32_t)s_txt[b + 1] << 8) |
              ((uint32_t)s_txt[b + 2] << 16) | ((uint32_t)s_txt[b + 3] << 24);
    }
    const int tail0 = head + 4 * nw;
    if (t < (int)rtotal - tail0) dst[tail0 + t] = s_txt[tail0 + t];
    g0 += rtotal;
#pragma unroll
    for (int k = 0; k < STRIDE; ++k) cur[k] = nxt[k];
  }
}

// ' ', '\t', '\n', '\v', '\f', '\r' (codes 32 and 9..13) as one mask test
__device__ inline bool is_space(uint8_t c) {
  return c <= 32 && ((0x100003E00ull >> c) & 1ull);
}

// Token parse from the staged bytes: an optional sign and digits, at most
// 10 of them significant (leading zeros are skipped, as operator>> skips
// them), ending at a space (bytes past the text read as spaces), within
// int32.  A token running past the staged bytes (only one with leading
// zeros can) continues from global memory: byte q of s is text[g0 + q].
__device__ inline bool parse_staged(const uint8_t *s, int p, int staged, const uint8_t *text,
                                    int64_t g0, int64_t len, int32_t *out) {
  auto at = [&](int q) -> uint8_t {
    if (q < staged) return s[q];
    return g0 + q < len ? text[g0 + q] : (uint8_t)' ';
  };
  bool neg = false;
  uint8_t c = at(p);
  if (c == '+' || c == '-') {
    neg = c == '-';
    ++p;
  }
  int64_t x = 0;
  int nd = 0;
  bool digit = false;
  for (;; ++p) {
    c = at(p);
    if (is_space(c)) break;
    if (c < '0' || c > '9') return false;
    digit = true;
    if (nd == 0 && c == '0') continue;
    if (nd >= 10) return false;
    x = x * 10 + (c - '0');
    ++nd;
  }
  if (!digit) return false;
  if (neg) x = -x;
  if (x < INT32_MIN || x > INT32_MAX) return false;
  *out = (int32_t)x;
  return true;
}

// A tile is kParseRounds rounds of kParseTile bytes with one look-back: the
// rounds' token-start counts are scanned together as 16-bit fields of one
// word (a round holds at most kParseTile / 2 starts), then each round's
// bytes and the 16 after it are staged in LDS (a token starting in the round
// without leading zeros is decided within 12 bytes: sign, 10 digits, one
// more byte; a longer one reads on from global memory), its tokens
// parsed into LDS in token order and written out as one run of records.
template <int DIM>
__global__ __launch_bounds__(kParseBlock) void parse_text_kernel(
    const uint8_t *__restrict__ text, int64_t len, int64_t N, int32_t *__restrict__ octree,
    uint64_t *__restrict__ ws, int ntiles) {
  constexpr int STRIDE = (1 << DIM) + 1;
  constexpr int kHalo = 16;
  __shared__ __align__(16) uint8_t s_txt[kParseTile + kHalo];
  __shared__ int32_t s_val[kParseTile / 2];
  __shared__ uint64_t s_w[kParseBlock / 64];
  __shared__ int s_tile;
  const int t = threadIdx.x;
  if (t == 0) s_tile = (int)atomicAdd(reinterpret_cast<unsigned int *>(ws + kWsCounter), 1u);
  __syncthreads();
  const int tile = s_tile;
  const int64_t tb = (int64_t)tile * kParseTile * kParseRounds;
  uint4 qs[kParseRounds];          // this thread's bytes of every round
  uint32_t smask[kParseRounds];    // and where tokens start in them
  uint8_t halo[kParseRounds];      // threads below kHalo: byte t after each round
  uint64_t packed = 0;
#pragma unroll
  for (int r = 0; r < kParseRounds; ++r) {
    const int64_t p0 = tb + (int64_t)r * kParseTile + (int64_t)t * kParseBytes;
    if (p0 + kParseBytes <= len && (((uintptr_t)(text + p0)) & 15) == 0) {
      qs[r] = *reinterpret_cast<const uint4 *>(text + p0);
    } else {
      uint32_t wv[4] = {0, 0, 0, 0};
#pragma unroll
      for (int k = 0; k < kParseBytes; ++k)
        wv[k >> 2] |= (uint32_t)(p0 + k < len ? text[p0 + k] : (uint8_t)' ') << (8 * (k & 3));
      qs[r] = make_uint4(wv[0], wv[1], wv[2], wv[3]);
    }
    bool prev_space = p0 == 0 || p0 > len || is_space(text[p0 - 1]);
    const uint32_t qw[4] = {qs[r].x, qs[r].y, qs[r].z, qs[r].w};
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < kParseBytes; ++k) {
      const bool sp = is_space((uint8_t)(qw[k >> 2] >> (8 * (k & 3))));
      if (!sp && prev_space) m |= 1u << k;
      prev_space = sp;
    }
    smask[r] = m;
    packed |= (uint64_t)__popc(m) << (16 * r);
    const int64_t ph = tb + (int64_t)(r + 1) * kParseTile + t;
    halo[r] = (t < kHalo && ph < len) ? text[ph] : (uint8_t)' ';
  }
  uint64_t tot_packed;
  const uint64_t ex = block_excl_sum<uint64_t, kParseBlock / 64>(packed, s_w, &tot_packed);
  uint32_t total = 0;
#pragma unroll
  for (int r = 0; r < kParseRounds; ++r) total += (uint32_t)(tot_packed >> (16 * r)) & 0xffffu;
  const uint64_t before = tile_prefix<kParseBlock>(ws + kWsStatus, tile, total);
  if (tile == ntiles - 1 && t == 0) ws[kWsTotal] = before + total;
  const uint64_t last = (uint64_t)N * STRIDE;   // tokens 1..last are the records
  uint64_t base = before;
#pragma unroll
  for (int r = 0; r < kParseRounds; ++r) {
    const uint32_t rtotal = (uint32_t)(tot_packed >> (16 * r)) & 0xffffu;
    if (r > 0) __syncthreads();   // the previous round is done with s_txt and s_val
    *reinterpret_cast<uint4 *>(s_txt + t * kParseBytes) = qs[r];
    if (t < kHalo) s_txt[kParseTile + t] = halo[r];
    __syncthreads();
    uint32_t starts = smask[r];
    uint32_t li = (uint32_t)(ex >> (16 * r)) & 0xffffu;
    while (starts) {
      const int j = __ffs(starts) - 1;
      starts &= starts - 1;
      const uint64_t k = base + li;
      if (k >= 1 && k <= last) {
        int32_t val = 0;
        if (!parse_staged(s_txt, t * kParseBytes + j, kParseTile + kHalo, text,
                          tb + (int64_t)r * kParseTile, len, &val))
          atomicMin(reinterpret_cast<unsigned long long *>(ws + kWsBad), (unsigned long long)k + 1);
        s_val[li] = val;
      }
      ++li;
    }
    __syncthreads();
    for (uint32_t i = t; i < rtotal; i += kParseBlock) {
      const uint64_t k = base + i;
      if (k >= 1 && k <= last) {
        const uint64_t rr = k - 1;
        const int32_t v = s_val[i];
        octree[rr] = (rr % STRIDE == STRIDE - 1) ? (int32_t)(uint8_t)v : v;
      }
    }
    base += rtotal;
  }
}

}  // namespace

int64_t text_ws_words(int64_t tiles) { return kWsStatus + tiles; }
int64_t format_text_tiles(int64_t O) {
  return (O + kFmtBlock * kFmtRounds - 1) / (kFmtBlock * kFmtRounds);
}
int64_t parse_text_tiles(int64_t len) {
  return (len + (int64_t)kParseTile * kParseRounds - 1) / ((int64_t)kParseTile * kParseRounds);
}

void launch_format_text(const int32_t *octree, int64_t O, int dim, int hdr_len, uint8_t *text,
                        int64_t capacity, uint64_t *ws, hipStream_t s) {
  const int64_t tiles = format_text_tiles(O);
  const int64_t words = text_ws_words(tiles);
  ws_zero_kernel<<<(unsigned)std::min<int64_t>((words + kBlock - 1) / kBlock, 1024), kBlock, 0,
                   s>>>(ws, words);
  if (dim == 2)
    format_text_kernel<2><<<(unsigned)tiles, kFmtBlock, 0, s>>>(octree, O, O, hdr_len, text,
                                                                capacity, ws, (int)tiles);
  else
    format_text_kernel<3><<<(unsigned)tiles, kFmtBlock, 0, s>>>(octree, O, O, hdr_len, text,
                                                                capacity, ws, (int)tiles);
}

void launch_parse_text(const uint8_t *text, int64_t len, int dim, int64_t N, int32_t *octree,
                       uint64_t *ws, hipStream_t s) {
  const int64_t tiles = parse_text_tiles(len);
  const int64_t words = text_ws_words(tiles);
  ws_zero_kernel<<<(unsigned)std::min<int64_t>((words + kBlock - 1) / kBlock, 1024), kBlock, 0,
                   s>>>(ws, words);
  if (dim == 2)
    parse_text_kernel<2><<<(unsigned)tiles, kParseBlock, 0, s>>>(text, len, N, octree, ws,
                                                                 (int)tiles);
  else
    parse_text_kernel<3><<<(unsigned)tiles, kParseBlock, 0, s>>>(text, len, N, octree, ws,
                                                                 (int)tiles);
}

}  // namespace pgvd
