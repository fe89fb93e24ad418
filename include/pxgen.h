/* Seeded synthetic event source for the pixel reconstruction chain.
 *
 * The reference reads recorded CMS events and conditions from binary files
 * (bin/Source.cc:9-21,50-73 for raw.bin; the ES producers for fedIds.bin,
 * cablingMap.bin, gain.bin, cpefast.bin, beamspot.bin).  Those files are not
 * available here, so this library builds a Phase-1-like pixel detector, its
 * conditions in the reference's binary layouts, and events of helix tracks
 * encoded as FED raw data, and can write them in the same file formats.
 *
 * A serialised event ("raw record") is exactly one event record of raw.bin:
 *   uint32 nfeds; nfeds x { uint32 fedId; uint32 nbytes; uint8 bytes[nbytes] }.
 */
#ifndef PXGEN_H
#define PXGEN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pxgen_detector pxgen_detector;
typedef struct pxgen_event pxgen_event;

/* Pile-up modes for pxgen_config.pu_mode */
enum { PXGEN_PU_FIXED = 0, PXGEN_PU_UNIFORM = 1, PXGEN_PU_POISSON = 2 };
/* Transverse-momentum spectra for pxgen_config.pt_mode */
enum { PXGEN_PT_LOGUNIFORM = 0, PXGEN_PT_INVSQUARE = 1 };

typedef struct {
  int32_t pu_mode;        /* number of vertices: fixed / uniform [pu_lo,pu_hi] / Poisson(pu_mean) clipped */
  int32_t pu_lo, pu_hi;   /* fixed mode uses pu_lo */
  double pu_mean;
  int32_t tracks_fixed;   /* >0: exactly this many tracks per vertex */
  double tracks_mean;     /* otherwise Poisson(tracks_mean) per vertex */
  int32_t pt_mode;
  double pt_lo, pt_hi;    /* GeV */
  double eta_max;
  double vz_sigma;        /* cm */
  double noise_frac;      /* probability per pixel of a noise hit */
  double noise_q_lo, noise_q_hi; /* noise charge range, electrons */
  double error_rate;      /* probability per FED of appended error words (tests) */
  uint64_t seed;          /* base seed; event i uses seed + i */
} pxgen_config;

/* Preset workloads.  0: 100-event serial sample (1 vertex x 30 tracks, no noise);
 * 1: PU50; 2: PU uniform 20-80; 3: PU Poisson(60) clipped 10-200; 4: PU200. */
int pxgen_preset(int which, pxgen_config* out);

pxgen_detector* pxgen_detector_create(uint64_t seed);
void pxgen_detector_destroy(pxgen_detector*);

/* Conditions, in the reference's in-memory / on-disk layouts. */
const void* pxgen_cabling(const pxgen_detector*);          /* SiPixelFedCablingMapGPU, 1440128 bytes */
const uint8_t* pxgen_mod_to_unp(const pxgen_detector*, uint32_t* n);
const void* pxgen_gain(const pxgen_detector*);             /* SiPixelGainForHLTonGPU, 24048 bytes */
const uint8_t* pxgen_gain_data(const pxgen_detector*, uint32_t* nbytes);
const void* pxgen_cpe_common(const pxgen_detector*);
const void* pxgen_cpe_det(const pxgen_detector*, uint32_t* n);
const void* pxgen_cpe_layer(const pxgen_detector*);
const void* pxgen_cpe_average(const pxgen_detector*);
const void* pxgen_beamspot(const pxgen_detector*);
const uint32_t* pxgen_fed_ids(const pxgen_detector*, uint32_t* n);
/* Writes fedIds.bin, cablingMap.bin, gain.bin, cpefast.bin, beamspot.bin into dir. */
int pxgen_write_conditions(const pxgen_detector*, const char* dir);

typedef struct {
  uint32_t n_vertices;
  uint32_t n_tracks;
  uint32_t n_crossings;   /* module crossings that left charge */
  uint32_t n_signal_pixels;
  uint32_t n_noise_pixels;
  uint32_t n_words;       /* pixel + error words over all FEDs */
  uint32_t n_error_words;
  float vertex_z[1024];   /* first 1024 vertex z positions, cm */
} pxgen_truth;

pxgen_event* pxgen_event_create(const pxgen_detector*, const pxgen_config*, uint32_t index);
void pxgen_event_destroy(pxgen_event*);
size_t pxgen_event_raw_size(const pxgen_event*);
const uint8_t* pxgen_event_raw(const pxgen_event*); /* the raw record described above */
void pxgen_event_truth(const pxgen_event*, pxgen_truth* out);

/* Appends n events to dir/raw.bin (created) after writing the conditions. */
int pxgen_write_dataset(const pxgen_detector*, const pxgen_config*, uint32_t n, const char* dir);

#ifdef __cplusplus
}
#endif
#endif
