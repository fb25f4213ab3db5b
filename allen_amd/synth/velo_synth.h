/* Synthetic VELO events in Allen's raw-bank layout, with Monte-Carlo truth.
 *
 * Events are straight-line tracks from Gaussian primary vertices, intersected with the
 * real sensor planes of velo_geometry.bin, digitised into 1-3 pixel clusters, plus
 * uniform noise pixels, and encoded as super-pixel (SP) raw banks:
 *   version 4: 104 banks per event, one per sensor pair, sourceID & 0x1FF = pair,
 *              SP word bit 31 = no 8-neighbour SP on the same sensor, bit 23 = sensor in pair,
 *              bits 22..14 = SP column, 13..8 = SP row, 7..0 = pixels (bit = 4*col + row);
 *   version 2: 208 banks per event, one per sensor, first word = number of SP words.
 * With `retina` set, every generated cluster is encoded instead as one retina cluster word
 * (bank type 73, the FPGA-clustered format decode_retinaclusters reads):
 *   version 4: 104 banks per event, one per sensor pair, bit 22 = sensor in pair;
 *   version 3: 208 banks per event, one per sensor, first word = number of cluster words;
 *   word bits 21..12 = column of the centroid, 11..10 = its quarter-pixel fraction,
 *   9..2 = row, 1..0 = its quarter-pixel fraction (centroids truncated to quarter pixels);
 *   identical words of a sensor are merged, and a bank with an odd number of words is padded
 *   with a zero word.
 * The event layout is the one main/src/Transpose.cpp:430-498 writes (see velo_amd.h). */
#ifndef VELO_SYNTH_H
#define VELO_SYNTH_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  unsigned n_events;
  uint64_t seed;
  int pileup;                    /* 0: Poisson(mean_tracks) tracks from one PV; 1: nu x NB tracks */
  double mean_tracks;            /* pileup == 0 */
  double nu;                     /* pileup == 1: collisions per event ~ Poisson(nu) */
  double nb_mean, nb_k;          /* tracks per collision ~ NegBinomial(mean, k) */
  double high_occupancy_fraction;/* fraction of events generated with nu = high_nu */
  double high_nu;
  double eta_min, eta_max;
  double backward_fraction;      /* fraction of tracks with eta in [-eta_max, -eta_min] */
  double pv_sigma_xy, pv_sigma_z;
  double hit_efficiency;
  double noise_fraction;         /* noise pixels per signal pixel */
  double dead_asic_fraction;
  double cluster_size_p[3];      /* probabilities of 1, 2 and 3 pixel clusters */
  int bank_version;              /* 4 or 2 (super pixels); 4 or 3 (retina) */
  int shuffle_sp;                /* shuffle SP words inside each bank */
  int retina;                    /* retina cluster banks instead of super-pixel banks */
} velo_synth_config;

typedef struct {
  uint32_t event;
  uint32_t pv;
  float x0, y0, z0, tx, ty, eta, phi;
  uint32_t n_hits;               /* digitised hits (clusters) */
  uint32_t n_modules;            /* distinct modules with a hit */
  uint32_t first_module, last_module;
  float p, pt;                   /* momentum and transverse momentum (MeV): pT ~ exponential of mean
                                    500 MeV, p = pT cosh(eta); drawn from a stream of their own, so
                                    they change nothing else of the event */
} velo_synth_particle;

typedef struct {
  uint32_t event;
  uint32_t particle;             /* index into the particle table of the batch, or 0xFFFFFFFF for noise */
  uint32_t lhcb_id;              /* LHCbID of the centroid pixel of the generated cluster (retina: with
                                    the fraction flags of the encoded centroid in bits 26-27) */
  uint16_t sensor;
  uint16_t n_pixels;
  float x, y, z;                 /* true intersection point */
} velo_synth_hit;

void velo_synth_default_config(velo_synth_config* c);
/* Generates c->n_events events; returns a handle or NULL on bad geometry. */
void* velo_synth_generate(const velo_synth_config* c, const char* geometry, size_t geometry_bytes, unsigned n_threads);
/* names: raw, offsets, sizes, types, particles, particle_offsets, hits, hit_offsets */
int velo_synth_get(void* h, const char* name, const void** ptr, size_t* nbytes);
void velo_synth_free(void* h);

#ifdef __cplusplus
}
#endif
#endif
