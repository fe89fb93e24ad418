// Host side of SiPixelRawToClusterCUDA: the FED loop that checks every pixel
// FED's CRC bit, header and trailer framing and concatenates the payload
// words for the decoder (plugin-SiPixelClusterizer/SiPixelRawToClusterCUDA.cc:
// 87-138, ErrorChecker.cc:38-91, SiPixelRawToClusterGPUKernel.cc:41-47).
//
// Per FED in the conditions' FED list (the pilot-blade FED 40 skipped):
//   * an empty buffer contributes nothing;
//   * a trailer with the CRC-modified bit (bit 2) set drops the FED;
//   * headers are consumed while they carry the 0x5 marker and the
//     "more headers" bit, trailers likewise from the end with the 0xA marker;
//   * the 32-bit words between the last header and the first trailer are
//     appended, and fedIds gets fedId - 1200 once per two words.
// Framing problems are reported through the diagnostics counters (the
// reference prints them; its checker records no error objects).
#include "fed_unpack.h"

#include <cstring>

namespace pxh {

  namespace {
    inline uint64_t word64(const uint8_t* p) {
      uint64_t v;
      std::memcpy(&v, p, 8);
      return v;
    }
  }  // namespace

  bool indexRawRecord(const uint8_t* raw, size_t len, std::vector<FedSpan>& out) {
    out.clear();
    if (len < 4)
      return false;
    uint32_t n;
    std::memcpy(&n, raw, 4);
    size_t pos = 4;
    for (uint32_t i = 0; i < n; ++i) {
      if (pos + 8 > len)
        return false;
      FedSpan f;
      std::memcpy(&f.fedId, raw + pos, 4);
      std::memcpy(&f.size, raw + pos + 4, 4);
      pos += 8;
      if (pos + f.size > len)
        return false;
      f.data = raw + pos;
      pos += f.size;
      out.push_back(f);
    }
    return pos == len;
  }

  int unpackFeds(const FedSpan* feds, size_t nSpans, const uint32_t* fedList, uint32_t nFedList, uint32_t* words,
                 uint8_t* fedIds, uint32_t capacity, uint32_t* nWordsOut, FedDiagnostics* diag) {
    FedDiagnostics d{};
    uint32_t count = 0;
    for (uint32_t k = 0; k < nFedList; ++k) {
      const uint32_t fedId = fedList[k];
      if (fedId == 40)
        continue;
      if (fedId < 1200)
        return -1;
      ++d.fedCounter;
      const FedSpan* span = nullptr;
      for (size_t j = 0; j < nSpans; ++j)
        if (feds[j].fedId == fedId) {
          span = &feds[j];
          break;
        }
      if (!span || span->size / 8 == 0)
        continue;
      const uint32_t n64 = span->size / 8;
      const uint8_t* base = span->data;
      if ((word64(base + 8 * size_t(n64 - 1)) >> 2) & 1) {
        ++d.crcErrors;
        continue;
      }
      size_t h = 0;  // index of the last header
      for (;;) {
        const uint64_t w = word64(base + 8 * h);
        const bool marker = (w >> 60) == 0x5;
        if (!marker) {
          break;
        }
        if (((w >> 8) & 0xFFF) != fedId)
          ++d.sourceIdMismatch;
        if (!((w >> 3) & 1))
          break;
        ++h;
      }
      size_t t = n64 - 1;  // index of the first trailer
      for (;;) {
        const uint64_t w = word64(base + 8 * t);
        const bool marker = (w >> 60) == 0xA;
        if (!marker) {
          ++d.trailerErrors;
          break;
        }
        if (((w >> 32) & 0xFFFFFF) != n64)
          ++d.lengthMismatch;
        if (!((w >> 3) & 1))
          break;
        --t;
      }
      const uint32_t length = uint32_t(2 * (t - h - 1));
      if (t < h + 1)
        continue;
      if (count + length > capacity)
        return -2;
      std::memcpy(words + count, base + 8 * (h + 1), 4 * size_t(length));
      std::memset(fedIds + count / 2, int(fedId - 1200), length / 2);
      count += length;
    }
    *nWordsOut = count;
    if (diag)
      *diag = d;
    return 0;
  }

}  // namespace pxh

extern "C" int pxh_unpack_raw_record(const uint8_t* raw, size_t len, const uint32_t* fedList, uint32_t nFedList,
                                     uint32_t* words, uint8_t* fedIds, uint32_t capacity, uint32_t* nWords) {
  std::vector<pxh::FedSpan> spans;
  if (!pxh::indexRawRecord(raw, len, spans))
    return -3;
  return pxh::unpackFeds(spans.data(), spans.size(), fedList, nFedList, words, fedIds, capacity, nWords, nullptr);
}
