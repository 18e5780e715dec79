// Content hash of a snapshot's edge part on the device.
//
// The reference hashes a snapshot with FNV-1a, byte by byte
// (src/common/graph.cpp:323-346): the header and coordinates, then per edge
// slot u and v (4 bytes each, little-endian), the weight (8) and the active
// byte.  A byte step is h <- (h xor b) * P mod 2^64 with P = 2^40 + 0x1b3.
// The xor touches only the low byte l of h: h xor b = h + d with
// d = (l xor b) - l, so h <- (h + d) * P, and the low byte evolves on its own,
// l <- ((l xor b) * 0xb3) mod 256.  A run of L bytes therefore maps h to
// h * P^L + G(l), where G depends only on the starting low byte, and maps
// the low byte through a permutation of the 256 values.
//
// The device evaluates the edge part from the host's prefix (header and
// coordinates, which never change during a run):
//   1. per chunk of E edge slots, the low-byte permutation for all 256
//      starting bytes;
//   2. per group of chunks, their composed permutation;
//   3. one walk over the groups from the prefix's low byte gives each
//      group's starting low byte, and 4. one walk per group each chunk's;
//   5. per chunk, G at its starting byte, weighted by P^(bytes after the
//      chunk), summed mod 2^64 with prefix * P^(all edge bytes).
// Every step is exact arithmetic mod 2^64 or mod 256, so the result equals
// the byte-serial hash.
#pragma once

#include "ee_runtime.h"

namespace ee {

class DeviceContentHash {
 public:
  // m edge slots (m > 0).
  void Allocate(size_t m, int device);
  // Enqueues on `stream` the hash of the edge part (per slot u, v, weight,
  // active != 0) continuing from `prefix`, written to *out (device memory).
  void Enqueue(uint64_t prefix, const int2* uv, const int64_t* weight, const int32_t* active, uint64_t* out,
               hipStream_t stream);
  size_t slots() const { return m_; }

 private:
  size_t m_ = 0;
  uint32_t edges_per_chunk_ = 0, group_size_ = 0;
  size_t chunks_ = 0, groups_ = 0;
  DeviceArray<uint8_t> perm_, group_perm_, chunk_start_, group_start_;
};

}  // namespace ee
