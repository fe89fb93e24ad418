// FED loop of the raw-to-cluster producer (see fed_unpack.cc).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace pxh {

  struct FedSpan {
    uint32_t fedId;
    uint32_t size;  // bytes
    const uint8_t* data;
  };

  struct FedDiagnostics {
    uint32_t fedCounter;
    uint32_t crcErrors;
    uint32_t sourceIdMismatch;
    uint32_t trailerErrors;
    uint32_t lengthMismatch;
  };

  // Splits one raw.bin event record into per-FED spans.
  bool indexRawRecord(const uint8_t* raw, size_t len, std::vector<FedSpan>& out);

  // Returns 0, -1 for a non-pixel FED id in the list, -2 when the words exceed capacity.
  int unpackFeds(const FedSpan* feds, size_t nSpans, const uint32_t* fedList, uint32_t nFedList, uint32_t* words,
                 uint8_t* fedIds, uint32_t capacity, uint32_t* nWords, FedDiagnostics* diag);

}  // namespace pxh

extern "C" int pxh_unpack_raw_record(const uint8_t* raw, size_t len, const uint32_t* fedList, uint32_t nFedList,
                                     uint32_t* words, uint8_t* fedIds, uint32_t capacity, uint32_t* nWords);
