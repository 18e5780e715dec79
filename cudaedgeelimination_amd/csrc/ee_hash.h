// Content hash of snapshots' edge parts on the device.
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
// coordinates, which never change during a run), for one or many snapshots
// at once (segments: consecutive ranges of one edge-slot array, as in the
// union of a fleet):
//   1. per chunk of E edge slots (never across segments), the low-byte
//      permutation for all 256 starting bytes;
//   2. per group of chunks, their composed permutation;
//   3. per segment, one walk over its groups from the prefix's low byte gives
//      each group's starting low byte (and the segment's result starts at
//      prefix * P^(all its edge bytes)); 4. one walk per group gives each
//      chunk's;
//   5. per chunk, G at its starting byte, weighted by P^(bytes after the
//      chunk in its segment), added to the segment's result mod 2^64.
// Every step is exact arithmetic mod 2^64 or mod 256, so each result equals
// the byte-serial hash.
#pragma once

#include <vector>

#include "ee_runtime.h"

namespace ee {

class DeviceContentHash {
 public:
  // Segment s owns edge slots [first[s], first[s + 1]); first is ascending,
  // with one more entry than there are segments.  node_offset[s] (or 0 for
  // every segment when empty) is subtracted from the stored endpoints before
  // hashing: a union of graphs stores endpoints as union node ids.  Uploads
  // the chunk, group and segment tables on `stream` and waits for them.
  void Allocate(const std::vector<size_t>& first, int device, hipStream_t stream,
                const std::vector<int32_t>& node_offset = {});
  // The segments' prefixes (host values, one per segment), uploaded on stream.
  void SetPrefixes(const uint64_t* prefixes, hipStream_t stream);
  // Enqueues on `stream` the hash of every segment's edge part (per slot u,
  // v, weight, active != 0) into out[s] (device memory, one per segment).
  void Enqueue(const int2* uv, const int64_t* weight, const int32_t* active, uint64_t* out, hipStream_t stream);
  size_t segments() const { return segments_; }

  struct Chunk {
    int64_t first, end, segment_end;  // edge slots
    int32_t segment, node_offset;
  };
  struct Group {
    uint32_t first_chunk, count, segment, pad;
  };
  struct Segment {
    uint32_t first_group, groups;
    uint64_t slots;
  };

 private:
  size_t segments_ = 0, chunks_ = 0, groups_ = 0;
  DeviceArray<Chunk> chunk_desc_;
  DeviceArray<Group> group_desc_;
  DeviceArray<Segment> segment_desc_;
  DeviceArray<uint64_t> prefixes_;
  DeviceArray<uint8_t> perm_, group_perm_, chunk_start_, group_start_;
};

}  // namespace ee
