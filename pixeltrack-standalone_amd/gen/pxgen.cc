// Seeded synthetic pixel detector and event generator (see include/pxgen.h).
//
// Geometry: a Phase-1-like pixel detector of 1856 modules of 160x416 pixels
// (100x150 um, rows 79/80 and the first/last column of every ROC double size).
//   barrel  4 layers, r = 2.9/6.8/10.9/16.0 cm, 12/28/44/64 ladders of 8
//           modules along z with a 6.7 cm pitch; ladder k is centred on
//           phi = 2*pi*(k+0.5)/n and even/odd ladders sit 0.15 cm in/out so
//           that neighbouring ladders overlap;
//   endcaps 3+3 disks at |z| = 32/39/48 cm, each an inner ring of 44 and an
//           outer ring of 68 radial modules, alternate modules 0.25 cm
//           before/after the disk plane (panel 1/2).
// Module indices follow the reference's layer ranges
// (Geometry/phase1PixelTopology.h:26-36) and barrel index = 8*ladder + module,
// which the CA "hole" checks rely on (plugin-PixelTriplets/GPUCACell.h:224-274).
// DetIds carry the bit fields the raw decoder reads
// (SiPixelRawToClusterGPUKernel.cc:404-418).
//
// Readout: module m is read by FED 1200 + m/18 on links 2*(m%18)+1 (ROCs 0-7
// of the module) and 2*(m%18)+2 (ROCs 8-15); ROC numbers inside a link run
// 1..8.  The pixel -> (ROC, ROC row/col) map is the inverse of the decoder's
// frame conversion (SiPixelRawToClusterGPUKernel.cc:72-138).
//
// Physics: straight vertices, helices in a uniform 3.8 T field, no material.
// Each module crossing deposits charge along the track path through the
// 285 um sensor, displaced by a Lorentz drift, shared among the pixels the
// path crosses; pixels under 1000 e- are dropped.  The charge per crossing is
// Moyal-distributed (Landau-like) with an MPV of 12000 e- per unit path, i.e.
// about 6000 e- per pixel for the common two-pixel cluster.  Charges are
// encoded to 8-bit ADC counts by inverting the gain/pedestal calibration.
#include "pxgen.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "../common/pxformats.h"

using namespace pxf;

namespace {

  // ---------------------------------------------------------------- random
  struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t& x) {
      uint64_t z = (x += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed, uint64_t stream = 0) {
      uint64_t x = seed * 0x2545F4914F6CDD1Dull + stream * 0x9E3779B97F4A7C15ull + 0x1234567ull;
      for (auto& v : s)
        v = splitmix(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {  // xoshiro256**
      uint64_t r = rotl(s[1] * 5, 7) * 9;
      uint64_t t = s[1] << 17;
      s[2] ^= s[0];
      s[3] ^= s[1];
      s[1] ^= s[2];
      s[0] ^= s[3];
      s[2] ^= t;
      s[3] = rotl(s[3], 45);
      return r;
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {
      double u1 = uniform();
      while (u1 <= 0)
        u1 = uniform();
      double u2 = uniform();
      return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
    int poisson(double mean) {
      if (mean <= 0)
        return 0;
      if (mean > 500) {
        int v = int(std::lround(mean + std::sqrt(mean) * normal()));
        return v < 0 ? 0 : v;
      }
      // multiply uniforms in log space (exact for any mean used here)
      double limit = -mean, acc = 0;
      int k = 0;
      while (true) {
        double u = uniform();
        while (u <= 0)
          u = uniform();
        acc += std::log(u);
        if (acc < limit)
          return k;
        ++k;
      }
    }
    // Moyal variate (mode 0): -ln(Z^2), Z standard normal
    double moyal() {
      double z = normal();
      while (z == 0)
        z = normal();
      return -std::log(z * z);
    }
  };

  struct V3 {
    double x, y, z;
  };
  inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
  inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
  inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
  inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
  inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

  // ---------------------------------------------------------------- geometry
  constexpr double kPitchX = 0.01, kPitchY = 0.015;
  constexpr double kThickness = 0.0285;
  constexpr double kHalfX = 0.5 * 162 * kPitchX;  // 0.81 cm (160 rows, 2 big)
  constexpr double kHalfY = 0.5 * 432 * kPitchY;  // 3.24 cm (416 cols, 16 big)
  constexpr double kBarrelR[4] = {2.9, 6.8, 10.9, 16.0};
  constexpr int kLadders[4] = {12, 28, 44, 64};
  constexpr double kModulePitchZ = 6.7;
  constexpr double kLadderStagger = 0.15;
  constexpr double kDiskZ[3] = {32.0, 39.0, 48.0};
  constexpr int kRingModules[2] = {44, 68};
  constexpr double kRingR[2] = {7.7, 12.9};
  constexpr double kPanelDz = 0.25;
  constexpr double kTanLorentzBarrel = 0.40, kTanLorentzEndcap = 0.08;
  constexpr double kBz = 3.8;                        // tesla
  constexpr double kCurvPerGeV = 0.0029979245 * kBz;  // 1/cm per (1/GeV)

  struct Module {
    int index;
    bool barrel;
    int layer;  // barrel layer 1-4, or disk 1-3
    int zside;  // endcap: +1 / -1; barrel: sign of z
    int ladder, modInLadder;
    int ring, posInRing, panel;
    V3 c, u, v, w;  // centre and local x, y, z axes
    uint32_t rawId;
    int fed, linkA, linkB;
    double lorentz;  // half the Lorentz displacement across the sensor, cm
  };

  struct CablingSlot {
    int module;
    int rocInDet;
  };

  // pixel row <-> local x units (100 um), with rows 79 and 80 double size
  inline int unitToRow(double ux) {  // ux in [0,162)
    int u = int(std::floor(ux));
    if (u < 79)
      return u;
    if (u < 81)
      return 79;
    if (u < 83)
      return 80;
    return u - 2;
  }
  // pixel col <-> local y units (150 um); in every ROC of 52 columns the first
  // and last are double size: ROC r spans units [54r, 54r+54)
  inline int unitToCol(double uy) {  // uy in [0,432)
    int u = int(std::floor(uy));
    int r = u / 54, k = u - 54 * r;
    int inRoc = k < 2 ? 0 : (k >= 52 ? 51 : k - 1);
    return 52 * r + inRoc;
  }

}  // namespace

struct pxgen_detector {
  std::vector<Module> modules;
  std::unique_ptr<SiPixelFedCablingMapGPU> cabling;
  std::vector<uint8_t> modToUnp;
  SiPixelGainForHLTonGPU gain;
  std::vector<uint8_t> gainData;
  CPECommonParams cpeCommon;
  std::vector<CPEDetParams> cpeDet;
  CPELayerGeometry cpeLayer;
  AverageGeometry avgGeom;
  BeamSpotPOD beamSpot;
  std::vector<uint32_t> fedIds;
  // per module calibration decoding, kept in double for the ADC inversion
  std::vector<uint8_t> deadOrNoisy;  // per (module, col, block)
};

struct pxgen_event {
  std::vector<uint8_t> raw;
  pxgen_truth truth;
};

namespace {

  void buildGeometry(pxgen_detector& d) {
    d.modules.resize(kNumModules);
    // barrel
    for (int L = 0; L < 4; ++L) {
      int n = kLadders[L];
      for (int k = 0; k < n; ++k) {
        double phi = 2 * M_PI * (k + 0.5) / n;
        double r = kBarrelR[L] + ((k % 2) ? kLadderStagger : -kLadderStagger);
        for (int m = 0; m < 8; ++m) {
          int idx = kLayerStart[L] + 8 * k + m;
          Module& md = d.modules[idx];
          md.index = idx;
          md.barrel = true;
          md.layer = L + 1;
          md.ladder = k;
          md.modInLadder = m;
          md.ring = md.posInRing = md.panel = 0;
          double zc = (m - 3.5) * kModulePitchZ;
          md.zside = zc < 0 ? -1 : 1;
          md.c = {r * std::cos(phi), r * std::sin(phi), zc};
          md.u = {-std::sin(phi), std::cos(phi), 0};
          md.v = {0, 0, 1};
          md.w = cross(md.u, md.v);
          md.rawId = (1u << 28) | (1u << 25) | (uint32_t(L + 1) << 20) | (uint32_t(k + 1) << 12) | (uint32_t(m + 1) << 2);
          md.lorentz = 0.5 * kTanLorentzBarrel * kThickness;
        }
      }
    }
    // endcaps: positive disks are layers 4-6, negative 7-9
    for (int side = 0; side < 2; ++side) {
      int zs = side == 0 ? 1 : -1;
      for (int D = 0; D < 3; ++D) {
        int base = kLayerStart[4 + 3 * side + D];
        int idx = base;
        for (int ring = 0; ring < 2; ++ring) {
          int n = kRingModules[ring];
          for (int k = 0; k < n; ++k, ++idx) {
            Module& md = d.modules[idx];
            double phi = 2 * M_PI * (k + 0.5) / n;
            md.index = idx;
            md.barrel = false;
            md.layer = D + 1;
            md.zside = zs;
            md.ladder = md.modInLadder = 0;
            md.ring = ring;
            md.posInRing = k;
            md.panel = 1 + (k % 2);
            double z = zs * (kDiskZ[D] + (md.panel == 1 ? -kPanelDz : kPanelDz));
            md.c = {kRingR[ring] * std::cos(phi), kRingR[ring] * std::sin(phi), z};
            md.v = {std::cos(phi), std::sin(phi), 0};
            md.u = zs > 0 ? V3{std::sin(phi), -std::cos(phi), 0} : V3{-std::sin(phi), std::cos(phi), 0};
            md.w = cross(md.u, md.v);
            int blade = (idx - base) / 2 + 1;
            md.rawId = (1u << 28) | (2u << 25) | (uint32_t(zs > 0 ? 2 : 1) << 23) | (uint32_t(D + 1) << 18) |
                       (uint32_t(blade) << 12) | (uint32_t(md.panel) << 10) | (1u << 2);
            md.lorentz = 0.5 * kTanLorentzEndcap * kThickness;
          }
        }
      }
    }
    for (auto& md : d.modules) {
      md.fed = kFirstPixelFed + md.index / 18;
      md.linkA = 2 * (md.index % 18) + 1;
      md.linkB = md.linkA + 1;
    }
  }

  uint32_t cablingIndex(int fed, int link, int roc) { return (fed - kFirstPixelFed) * kMaxLink * kMaxRoc + (link - 1) * kMaxRoc + roc; }

  void buildCabling(pxgen_detector& d, Rng& rng) {
    d.cabling = std::make_unique<SiPixelFedCablingMapGPU>();
    auto& c = *d.cabling;
    std::memset(&c, 0, sizeof(c));
    for (uint32_t i = 0; i < kCablingSize; ++i) {
      c.RawId[i] = 9999;
      c.rocInDet[i] = 9999;
      c.moduleId[i] = 9999;
    }
    uint32_t maxIndex = 0;
    for (auto& md : d.modules) {
      for (int rid = 0; rid < 16; ++rid) {
        int link = rid < 8 ? md.linkA : md.linkB;
        int roc = rid % 8 + 1;
        uint32_t i = cablingIndex(md.fed, link, roc);
        c.fed[i] = md.fed;
        c.link[i] = link;
        c.roc[i] = roc;
        c.RawId[i] = md.rawId;
        c.rocInDet[i] = rid;
        c.moduleId[i] = md.index;
        maxIndex = std::max(maxIndex, i);
      }
    }
    c.size = maxIndex + 1;
    // a handful of bad ROCs
    for (int k = 0; k < 8; ++k) {
      auto& md = d.modules[size_t(rng.uniform() * kNumModules)];
      int rid = int(rng.uniform() * 16);
      c.badRocs[cablingIndex(md.fed, rid < 8 ? md.linkA : md.linkB, rid % 8 + 1)] = 1;
    }
    d.modToUnp.assign(kCablingSize, 0);
    for (auto& md : d.modules)
      if (std::find(d.fedIds.begin(), d.fedIds.end(), uint32_t(md.fed)) == d.fedIds.end())
        d.fedIds.push_back(md.fed);
    std::sort(d.fedIds.begin(), d.fedIds.end());
  }

  constexpr uint32_t kColBytes = 4;  // two 80-row blocks x {gain, ped}
  constexpr uint32_t kModuleGainBytes = kColsPerModule * kColBytes;

  void buildGains(pxgen_detector& d, Rng& rng) {
    auto& g = d.gain;
    std::memset(&g, 0, sizeof(g));
    g.minPed_ = 0.f;
    g.maxPed_ = 50.f;
    g.minGain_ = 2.0f;
    g.maxGain_ = 4.0f;
    g.nBinsToUseForEncoding_ = 253;
    g.pedPrecision = (g.maxPed_ - g.minPed_) / float(g.nBinsToUseForEncoding_);
    g.gainPrecision = (g.maxGain_ - g.minGain_) / float(g.nBinsToUseForEncoding_);
    g.numberOfRowsAveragedOver_ = 80;
    g.deadFlag_ = 255;
    g.noisyFlag_ = 254;
    for (uint32_t m = 0; m < kNumModules; ++m)
      g.rangeAndCols[m] = {m * kModuleGainBytes, (m + 1) * kModuleGainBytes, int32_t(kColsPerModule)};
    d.gainData.resize(size_t(kNumModules) * kModuleGainBytes);
    for (size_t i = 0; i < d.gainData.size(); i += 2) {
      double gv = 3.0 + 0.3 * rng.normal();
      double pv = 20.0 + 3.0 * rng.normal();
      int gb = int(std::lround((gv - g.minGain_) / g.gainPrecision));
      int pb = int(std::lround((pv - g.minPed_) / g.pedPrecision));
      d.gainData[i] = uint8_t(std::clamp(gb, 0, 253));
      d.gainData[i + 1] = uint8_t(std::clamp(pb, 0, 253));
    }
    // dead and noisy column blocks
    for (int k = 0; k < 128; ++k) {
      size_t blk = size_t(rng.uniform() * (d.gainData.size() / 2));
      d.gainData[2 * blk + 1] = uint8_t(k % 2 ? g.noisyFlag_ : g.deadFlag_);
    }
  }

  void buildCPE(pxgen_detector& d) {
    d.cpeCommon = {float(kThickness), float(kThickness), float(kPitchX), float(kPitchY)};
    d.cpeDet.resize(kNumModules);
    const BeamSpotPOD& bs = d.beamSpot;
    V3 origin{bs.x, bs.y, bs.z};
    for (auto& md : d.modules) {
      CPEDetParams& p = d.cpeDet[md.index];
      std::memset(&p, 0, sizeof(p));
      p.isBarrel = md.barrel;
      p.isPosZ = md.c.z > 0;
      p.layer = uint16_t(md.layer);
      p.index = uint16_t(md.index);
      p.rawId = md.rawId;
      p.shiftX = float(-md.lorentz);
      p.shiftY = 0.f;
      p.chargeWidthX = float(2 * md.lorentz);
      p.chargeWidthY = 0.f;
      V3 o = origin - md.c;
      p.x0 = float(dot(o, md.u));
      p.y0 = float(dot(o, md.v));
      p.z0 = float(dot(o, md.w));
      const float bsx[3] = {0.0018f, 0.0013f, 0.0020f}, bsy[3] = {0.0030f, 0.0021f, 0.0032f};
      const float esx[3] = {0.0020f, 0.0015f, 0.0022f}, esy[3] = {0.0028f, 0.0024f, 0.0030f};
      for (int i = 0; i < 3; ++i) {
        p.sx[i] = md.barrel ? bsx[i] : esx[i];
        p.sy[i] = md.barrel ? bsy[i] : esy[i];
      }
      auto& f = p.frame;
      f.px = float(md.c.x);
      f.py = float(md.c.y);
      f.pz = float(md.c.z);
      f.R11 = float(md.u.x), f.R12 = float(md.u.y), f.R13 = float(md.u.z);
      f.R21 = float(md.v.x), f.R22 = float(md.v.y), f.R23 = float(md.v.z);
      f.R31 = float(md.w.x), f.R32 = float(md.w.y), f.R33 = float(md.w.z);
    }
    for (uint32_t i = 0; i <= kNumLayers; ++i)
      d.cpeLayer.layerStart[i] = kLayerStart[i];
    for (uint32_t j = 0; j < kLayerIndexSize; ++j) {
      uint32_t m = j * kModuleStride;
      uint8_t l = 0;
      while (l < kNumLayers && m >= kLayerStart[l + 1])
        ++l;
      d.cpeLayer.layer[j] = l;
    }
    auto& ag = d.avgGeom;
    for (uint32_t il = 0; il < kLaddersInBarrel; ++il) {
      double sx = 0, sy = 0, sz = 0;
      for (int m = 0; m < 8; ++m) {
        auto& md = d.modules[8 * il + m];
        sx += md.c.x;
        sy += md.c.y;
        sz += md.c.z;
      }
      ag.ladderX[il] = float(sx / 8);
      ag.ladderY[il] = float(sy / 8);
      ag.ladderZ[il] = float(sz / 8);
      ag.ladderR[il] = float(std::hypot(sx / 8, sy / 8));
      ag.ladderMinZ[il] = float(sz / 8 - 4 * kModulePitchZ);
      ag.ladderMaxZ[il] = float(sz / 8 + 4 * kModulePitchZ);
    }
    ag.endCapZ[0] = float(kDiskZ[0] - kPanelDz);
    ag.endCapZ[1] = float(-(kDiskZ[0] - kPanelDz));
  }

  // ---------------------------------------------------------------- tracking
  struct Helix {
    V3 v;         // vertex
    double phi0;  // initial transverse direction
    double kappa; // signed curvature, 1/cm (positive = counter-clockwise)
    double cot;   // dz/ds
    V3 at(double s) const {
      if (std::abs(kappa) < 1e-12)
        return {v.x + s * std::cos(phi0), v.y + s * std::sin(phi0), v.z + s * cot};
      double a = phi0 + kappa * s;
      return {v.x + (std::sin(a) - std::sin(phi0)) / kappa, v.y - (std::cos(a) - std::cos(phi0)) / kappa, v.z + s * cot};
    }
    V3 dir(double s) const {  // d/ds, not normalised
      double a = phi0 + kappa * s;
      return {std::cos(a), std::sin(a), cot};
    }
  };

  // Newton solve of the helix crossing the module plane, starting near s0.
  bool intersect(const Helix& h, const Module& md, double s0, double& s, double& lx, double& ly) {
    s = s0;
    for (int it = 0; it < 20; ++it) {
      double g = dot(h.at(s) - md.c, md.w);
      double gp = dot(h.dir(s), md.w);
      if (std::abs(gp) < 1e-9)
        return false;
      double ds = g / gp;
      s -= ds;
      if (std::abs(ds) < 1e-10)
        break;
    }
    if (s <= 0 || std::abs(s - s0) > 3.0)
      return false;
    V3 p = h.at(s) - md.c;
    if (std::abs(dot(p, md.w)) > 1e-6)
      return false;
    lx = dot(p, md.u);
    ly = dot(p, md.v);
    return std::abs(lx) < kHalfX && std::abs(ly) < kHalfY;
  }

  using PixelMap = std::unordered_map<uint64_t, double>;  // key: module*2^17 + row*512 + col

  inline uint64_t pixKey(int m, int row, int col) { return (uint64_t(m) << 17) | (uint64_t(row) << 9) | uint64_t(col); }

  // Deposit charge along the path through the sensor.
  int deposit(const Helix& h, const Module& md, double s, double lx, double ly, Rng& rng, PixelMap& pix) {
    V3 t = h.dir(s);
    double tn = std::sqrt(dot(t, t));
    V3 dl{dot(t, md.u) / tn, dot(t, md.v) / tn, dot(t, md.w) / tn};
    if (std::abs(dl.z) < 1e-3)
      return 0;
    double dxdz = dl.x / dl.z, dydz = dl.y / dl.z;
    // entry (local z = -t/2) and exit (+t/2); drift toward +z adds up to 2*lorentz in x
    double x1 = lx - 0.5 * kThickness * dxdz + 2 * md.lorentz, y1 = ly - 0.5 * kThickness * dydz;
    double x2 = lx + 0.5 * kThickness * dxdz, y2 = ly + 0.5 * kThickness * dydz;
    double pathFactor = 1.0 / std::abs(dl.z);
    double q = pathFactor * (12000.0 + 1000.0 * rng.moyal());
    q = std::clamp(q, 3000.0 * pathFactor, 200000.0);
    double lenPix = std::abs(x2 - x1) / kPitchX + std::abs(y2 - y1) / kPitchY;
    int n = std::max(16, int(8 * lenPix));
    std::map<std::pair<int, int>, double> local;
    for (int i = 0; i < n; ++i) {
      double f = (i + 0.5) / n;
      double x = x1 + f * (x2 - x1), y = y1 + f * (y2 - y1);
      double ux = x / kPitchX + 81.0, uy = y / kPitchY + 216.0;
      if (ux < 0 || ux >= 162 || uy < 0 || uy >= 432)
        continue;
      local[{unitToRow(ux), unitToCol(uy)}] += q / n;
    }
    int npix = 0;
    for (auto& [rc, qq] : local) {
      if (qq < 1000.0)
        continue;
      pix[pixKey(md.index, rc.first, rc.second)] += qq;
      ++npix;
    }
    return npix;
  }

  // Inverse of the decoder's frame conversion: module pixel -> (ROC in module, ROC row, ROC col)
  void toRoc(const Module& md, int row, int col, int& rid, int& lrow, int& lcol) {
    bool typeA = md.barrel ? (md.modInLadder + 1 < 5 && md.layer != 1) : true;
    if (typeA) {
      if (row < 80) {
        rid = 7 - col / 52;
        lrow = row;
        lcol = (8 - rid) * 52 - 1 - col;
      } else {
        rid = 8 + col / 52;
        lrow = 159 - row;
        lcol = col - (rid - 8) * 52;
      }
    } else {
      if (row >= 80) {
        rid = col / 52;
        lrow = 159 - row;
        lcol = col - rid * 52;
      } else {
        rid = 15 - col / 52;
        lrow = row;
        lcol = (16 - rid) * 52 - 1 - col;
      }
    }
  }

  uint8_t encodeAdc(const pxgen_detector& d, int m, int row, int col, double q) {
    const auto& g = d.gain;
    uint32_t off = g.rangeAndCols[m].first + col * kColBytes + 2 * (row / 80);
    uint8_t gb = d.gainData[off], pb = d.gainData[off + 1];
    if (pb >= 254)  // dead/noisy: any value, the calibration drops the pixel
      pb = 0;
    double gain = gb * double(g.gainPrecision) + g.minGain_;
    double ped = pb * double(g.pedPrecision) + g.minPed_;
    bool l1 = m < 96;
    double conv = l1 ? 50.0 : 47.0, offs = l1 ? -670.0 : -60.0;
    double adc = (q - offs) / conv / gain + ped;
    return uint8_t(std::clamp(int(std::lround(adc)), 0, 255));
  }

  void appendU32(std::vector<uint8_t>& b, uint32_t v) {
    uint8_t t[4];
    std::memcpy(t, &v, 4);
    b.insert(b.end(), t, t + 4);
  }

}  // namespace

extern "C" {

int pxgen_preset(int which, pxgen_config* c) {
  if (!c)
    return -1;
  std::memset(c, 0, sizeof(*c));
  c->eta_max = 2.5;
  c->noise_q_lo = 500;
  c->noise_q_hi = 5000;
  c->seed = 20261202;
  c->pt_mode = PXGEN_PT_INVSQUARE;
  c->pt_lo = 0.5;
  c->pt_hi = 100.0;
  c->tracks_mean = 35;
  c->vz_sigma = 3.5;
  c->noise_frac = 0.00018;  // reproduces ~50k digis and ~12k clusters at PU50
  switch (which) {
    case 0:
      c->pu_mode = PXGEN_PU_FIXED;
      c->pu_lo = c->pu_hi = 1;
      c->tracks_fixed = 30;
      c->pt_mode = PXGEN_PT_LOGUNIFORM;
      c->pt_lo = 1.0;
      c->pt_hi = 10.0;
      c->vz_sigma = 5.0;
      c->noise_frac = 0;
      return 0;
    case 1:
      c->pu_mode = PXGEN_PU_FIXED;
      c->pu_lo = c->pu_hi = 50;
      return 0;
    case 2:
      c->pu_mode = PXGEN_PU_UNIFORM;
      c->pu_lo = 20;
      c->pu_hi = 80;
      return 0;
    case 3:
      c->pu_mode = PXGEN_PU_POISSON;
      c->pu_mean = 60;
      c->pu_lo = 10;
      c->pu_hi = 200;
      return 0;
    case 4:
      c->pu_mode = PXGEN_PU_FIXED;
      c->pu_lo = c->pu_hi = 200;
      return 0;
    default:
      return -1;
  }
}

pxgen_detector* pxgen_detector_create(uint64_t seed) {
  auto d = std::make_unique<pxgen_detector>();
  Rng rng(seed, 1);
  d->beamSpot = {0.05f, -0.03f, 0.2f, 3.5f, 0.0015f, 0.0015f, 0.f, 0.f, 0.f, 0.f, 0.f};
  buildGeometry(*d);
  buildCabling(*d, rng);
  buildGains(*d, rng);
  buildCPE(*d);
  return d.release();
}

void pxgen_detector_destroy(pxgen_detector* d) { delete d; }

const void* pxgen_cabling(const pxgen_detector* d) { return d->cabling.get(); }
const uint8_t* pxgen_mod_to_unp(const pxgen_detector* d, uint32_t* n) {
  if (n)
    *n = uint32_t(d->modToUnp.size());
  return d->modToUnp.data();
}
const void* pxgen_gain(const pxgen_detector* d) { return &d->gain; }
const uint8_t* pxgen_gain_data(const pxgen_detector* d, uint32_t* n) {
  if (n)
    *n = uint32_t(d->gainData.size());
  return d->gainData.data();
}
const void* pxgen_cpe_common(const pxgen_detector* d) { return &d->cpeCommon; }
const void* pxgen_cpe_det(const pxgen_detector* d, uint32_t* n) {
  if (n)
    *n = uint32_t(d->cpeDet.size());
  return d->cpeDet.data();
}
const void* pxgen_cpe_layer(const pxgen_detector* d) { return &d->cpeLayer; }
const void* pxgen_cpe_average(const pxgen_detector* d) { return &d->avgGeom; }
const void* pxgen_beamspot(const pxgen_detector* d) { return &d->beamSpot; }
const uint32_t* pxgen_fed_ids(const pxgen_detector* d, uint32_t* n) {
  if (n)
    *n = uint32_t(d->fedIds.size());
  return d->fedIds.data();
}

int pxgen_write_conditions(const pxgen_detector* d, const char* dir) {
  std::string base(dir);
  auto open = [&](const char* name) { return std::fopen((base + "/" + name).c_str(), "wb"); };
  FILE* f = open("fedIds.bin");
  if (!f)
    return -1;
  uint32_t n = uint32_t(d->fedIds.size());
  std::fwrite(&n, 4, 1, f);
  std::fwrite(d->fedIds.data(), 4, n, f);
  std::fclose(f);
  if (!(f = open("cablingMap.bin")))
    return -1;
  std::fwrite(d->cabling.get(), sizeof(SiPixelFedCablingMapGPU), 1, f);
  n = uint32_t(d->modToUnp.size());
  std::fwrite(&n, 4, 1, f);
  std::fwrite(d->modToUnp.data(), 1, n, f);
  std::fclose(f);
  if (!(f = open("gain.bin")))
    return -1;
  std::fwrite(&d->gain, sizeof(d->gain), 1, f);
  n = uint32_t(d->gainData.size());
  std::fwrite(&n, 4, 1, f);
  std::fwrite(d->gainData.data(), 1, n, f);
  std::fclose(f);
  if (!(f = open("cpefast.bin")))
    return -1;
  std::fwrite(&d->cpeCommon, sizeof(d->cpeCommon), 1, f);
  n = uint32_t(d->cpeDet.size());
  std::fwrite(&n, 4, 1, f);
  std::fwrite(d->cpeDet.data(), sizeof(CPEDetParams), n, f);
  std::fwrite(&d->avgGeom, sizeof(d->avgGeom), 1, f);
  std::fwrite(&d->cpeLayer, sizeof(d->cpeLayer), 1, f);
  std::fclose(f);
  if (!(f = open("beamspot.bin")))
    return -1;
  std::fwrite(&d->beamSpot, sizeof(d->beamSpot), 1, f);
  std::fclose(f);
  return 0;
}

pxgen_event* pxgen_event_create(const pxgen_detector* d, const pxgen_config* cfg, uint32_t index) {
  auto ev = std::make_unique<pxgen_event>();
  std::memset(&ev->truth, 0, sizeof(ev->truth));
  const uint64_t seed = cfg->seed + index;
  Rng rngV(seed, 11), rngT(seed, 12), rngQ(seed, 13), rngN(seed, 14), rngE(seed, 15);

  int nv = cfg->pu_lo;
  if (cfg->pu_mode == PXGEN_PU_UNIFORM)
    nv = cfg->pu_lo + int(rngV.uniform() * (cfg->pu_hi - cfg->pu_lo + 1));
  else if (cfg->pu_mode == PXGEN_PU_POISSON)
    nv = std::clamp(rngV.poisson(cfg->pu_mean), cfg->pu_lo, cfg->pu_hi);
  ev->truth.n_vertices = nv;

  PixelMap pix;
  pix.reserve(1 << 16);
  const auto& bs = d->beamSpot;
  for (int iv = 0; iv < nv; ++iv) {
    V3 vtx{bs.x + bs.beamWidthX * rngV.normal(), bs.y + bs.beamWidthY * rngV.normal(), bs.z + cfg->vz_sigma * rngV.normal()};
    if (iv < 1024)
      ev->truth.vertex_z[iv] = float(vtx.z);
    int nt = cfg->tracks_fixed > 0 ? cfg->tracks_fixed : rngV.poisson(cfg->tracks_mean);
    for (int it = 0; it < nt; ++it) {
      double pt;
      if (cfg->pt_mode == PXGEN_PT_LOGUNIFORM)
        pt = cfg->pt_lo * std::exp(rngT.uniform() * std::log(cfg->pt_hi / cfg->pt_lo));
      else {
        double u = rngT.uniform();
        pt = cfg->pt_lo / std::max(1.0 - u, cfg->pt_lo / cfg->pt_hi);
      }
      double eta = rngT.uniform(-cfg->eta_max, cfg->eta_max);
      double phi = rngT.uniform(-M_PI, M_PI);
      int q = rngT.uniform() < 0.5 ? -1 : 1;
      Helix h{vtx, phi, -q * kCurvPerGeV / pt, std::sinh(eta)};
      ev->truth.n_tracks++;
      // march along the first half turn, find layer/disk crossings
      double smax = std::min(M_PI / std::abs(h.kappa), 120.0);
      const double ds = 0.05;
      V3 prev = h.at(0);
      double rprev = std::hypot(prev.x, prev.y);
      for (double s = ds; s <= smax; s += ds) {
        V3 p = h.at(s);
        double r = std::hypot(p.x, p.y);
        if (r > 20 || std::abs(p.z) > 55)
          break;
        // barrel layers
        for (int L = 0; L < 4; ++L) {
          double rl = kBarrelR[L];
          if ((rprev - rl) * (r - rl) > 0)
            continue;
          if (std::abs(p.z) > 27.5)
            continue;
          int n = kLadders[L];
          double ph = std::atan2(p.y, p.x);
          if (ph < 0)
            ph += 2 * M_PI;
          int k0 = int(ph / (2 * M_PI / n));
          int m0 = int(std::floor((p.z + 4 * kModulePitchZ) / kModulePitchZ));
          for (int dk = -1; dk <= 1; ++dk)
            for (int dm = -1; dm <= 1; ++dm) {
              int k = ((k0 + dk) % n + n) % n, m = m0 + dm;
              if (m < 0 || m > 7)
                continue;
              const Module& md = d->modules[kLayerStart[L] + 8 * k + m];
              double sh, lx, ly;
              if (intersect(h, md, s, sh, lx, ly)) {
                int np = deposit(h, md, sh, lx, ly, rngQ, pix);
                if (np) {
                  ev->truth.n_crossings++;
                }
              }
            }
        }
        // disks
        for (int D = 0; D < 3; ++D)
          for (int zs = -1; zs <= 1; zs += 2) {
            double zd = zs * kDiskZ[D];
            if ((prev.z - zd) * (p.z - zd) > 0)
              continue;
            if (r < 4.0 || r > 16.6)
              continue;
            double ph = std::atan2(p.y, p.x);
            if (ph < 0)
              ph += 2 * M_PI;
            int base = kLayerStart[(zs > 0 ? 4 : 7) + D];
            for (int ring = 0; ring < 2; ++ring) {
              int n = kRingModules[ring];
              int k0 = int(ph / (2 * M_PI / n));
              for (int dk = -1; dk <= 1; ++dk) {
                int k = ((k0 + dk) % n + n) % n;
                const Module& md = d->modules[base + (ring ? kRingModules[0] : 0) + k];
                double sh, lx, ly;
                if (intersect(h, md, s, sh, lx, ly)) {
                  if (deposit(h, md, sh, lx, ly, rngQ, pix))
                    ev->truth.n_crossings++;
                }
              }
            }
          }
        prev = p;
        rprev = r;
      }
    }
  }
  ev->truth.n_signal_pixels = uint32_t(pix.size());
  // noise
  if (cfg->noise_frac > 0) {
    const double perModule = cfg->noise_frac * kRowsPerModule * kColsPerModule;
    for (uint32_t m = 0; m < kNumModules; ++m) {
      int nn = rngN.poisson(perModule);
      for (int i = 0; i < nn; ++i) {
        int row = int(rngN.uniform() * kRowsPerModule), col = int(rngN.uniform() * kColsPerModule);
        double q = rngN.uniform(cfg->noise_q_lo, cfg->noise_q_hi);
        auto key = pixKey(m, row, col);
        if (!pix.count(key))
          ev->truth.n_noise_pixels++;
        pix[key] += q;
      }
    }
  }
  // words per module, ordered by (link, roc, column/row in ROC)
  std::vector<std::vector<std::pair<uint64_t, uint32_t>>> perModule(kNumModules);
  for (auto& [key, q] : pix) {
    int m = int(key >> 17), row = int((key >> 9) & 0xFF), col = int(key & 0x1FF);
    const Module& md = d->modules[m];
    int rid, lrow, lcol;
    toRoc(md, row, col, rid, lrow, lcol);
    int link = rid < 8 ? md.linkA : md.linkB;
    int roc = rid % 8 + 1;
    uint32_t adc = encodeAdc(*d, m, row, col, q);
    uint32_t w = (uint32_t(link) << kLinkShift) | (uint32_t(roc) << kRocShift) | adc;
    uint64_t order;
    if (md.barrel && md.layer == 1) {
      w |= (uint32_t(lcol) << kColShiftL1) | (uint32_t(lrow) << kRowShiftL1);
      order = (uint64_t(link) << 40) | (uint64_t(roc) << 32) | (uint64_t(lcol) << 16) | uint64_t(lrow);
    } else {
      uint32_t dcol = lcol / 2, pxid = 2 * (80 - lrow) + (lcol % 2);
      w |= (dcol << kDcolShift) | (pxid << kPxidShift);
      order = (uint64_t(link) << 40) | (uint64_t(roc) << 32) | (uint64_t(dcol) << 16) | uint64_t(pxid);
    }
    perModule[m].push_back({order, w});
  }
  // FED payloads
  std::map<uint32_t, std::vector<uint32_t>> fedWords;
  for (uint32_t f : d->fedIds)
    fedWords[f];
  for (uint32_t m = 0; m < kNumModules; ++m) {
    auto& v = perModule[m];
    std::sort(v.begin(), v.end());
    auto& fw = fedWords[d->modules[m].fed];
    for (auto& [o, w] : v)
      fw.push_back(w);
  }
  if (cfg->error_rate > 0) {
    for (auto& [fed, fw] : fedWords) {
      if (rngE.uniform() >= cfg->error_rate)
        continue;
      // one module of this FED
      int m = int((fed - kFirstPixelFed) * 18 + int(rngE.uniform() * 18));
      if (m >= int(kNumModules))
        m = kNumModules - 1;
      const Module& md = d->modules[m];
      uint32_t link = md.linkA;
      const uint32_t types[] = {25, 26, 27, 28, 29, 30, 31};
      for (uint32_t t : types) {
        uint32_t w = (link << kLinkShift) | (t << kRocShift);
        if (t == 29)
          w |= (uint32_t(1 + int(rngE.uniform() * 4)) << 8) | 0x3;  // block + channel bits
        if (t == 30)
          w |= (rngE.uniform() < 0.5 ? 1u : 8u) << 8;  // state-match bits
        fw.push_back(w);
        ev->truth.n_error_words++;
      }
      // an invalid pixel address on this module
      if (md.barrel && md.layer == 1)
        fw.push_back((link << kLinkShift) | (1u << kRocShift) | (90u << kRowShiftL1) | (3u << kColShiftL1) | 50u);
      else
        fw.push_back((link << kLinkShift) | (1u << kRocShift) | (30u << kDcolShift) | (7u << kPxidShift) | 50u);
      ev->truth.n_error_words++;
    }
  }
  // serialise: nfeds, then {fedId, size, bytes}; each FED = header + payload + trailer
  auto& raw = ev->raw;
  appendU32(raw, uint32_t(fedWords.size()));
  uint32_t lvl1 = index + 1;
  for (auto& [fed, fw] : fedWords) {
    if (fw.size() % 2)
      fw.push_back(0);  // pad to whole 64-bit words
    ev->truth.n_words += uint32_t(fw.size());
    uint32_t n64 = uint32_t(fw.size() / 2 + 2);
    appendU32(raw, fed);
    appendU32(raw, n64 * 8);
    // header: low word = source id word, high word = event id word with marker 0x5
    appendU32(raw, (fed & 0xFFFu) << 8);
    appendU32(raw, (0x5u << 28) | (1u << 24) | (lvl1 & 0xFFFFFFu));
    for (uint32_t w : fw)
      appendU32(raw, w);
    // trailer: low word = consistency check (CRC bit 2 clear), high = marker 0xA + length
    appendU32(raw, 0u);
    appendU32(raw, (0xAu << 28) | (n64 & 0xFFFFFFu));
  }
  return ev.release();
}

void pxgen_event_destroy(pxgen_event* e) { delete e; }
size_t pxgen_event_raw_size(const pxgen_event* e) { return e->raw.size(); }
const uint8_t* pxgen_event_raw(const pxgen_event* e) { return e->raw.data(); }
void pxgen_event_truth(const pxgen_event* e, pxgen_truth* out) { *out = e->truth; }

int pxgen_write_dataset(const pxgen_detector* d, const pxgen_config* cfg, uint32_t n, const char* dir) {
  if (pxgen_write_conditions(d, dir))
    return -1;
  FILE* f = std::fopen((std::string(dir) + "/raw.bin").c_str(), "wb");
  if (!f)
    return -1;
  for (uint32_t i = 0; i < n; ++i) {
    pxgen_event* e = pxgen_event_create(d, cfg, i);
    std::fwrite(e->raw.data(), 1, e->raw.size(), f);
    pxgen_event_destroy(e);
  }
  std::fclose(f);
  return 0;
}

}  // extern "C"
