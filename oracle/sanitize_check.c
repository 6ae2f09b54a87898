/* sanitize_check.c — TEST INFRASTRUCTURE ONLY: drives every oracle entry
 * point over one .csdb fixture so that a build with AddressSanitizer and
 * UndefinedBehaviorSanitizer (host code) checks the oracle's memory safety:
 * parse, rANS of every brick, every brick at every LOD, the threaded
 * pyramid decode, the first-hit and compositing renders (orthographic and
 * perspective), the LOD rule, plus out-of-range brick indices and a
 * truncated buffer. Usage: sanitize_check <file.csdb> <brick_size>.
 * Prints "ok" and exits 0 when every call returned as expected. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static uint64_t pyr_off(uint32_t t) { return (((uint64_t)1 << (3 * t)) - 1) / 7; }

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  fseek(f, 0, SEEK_END);
  const long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  uint8_t* buf = (uint8_t*)malloc((size_t)n);
  if (fread(buf, 1, (size_t)n, f) != (size_t)n) return 2;
  fclose(f);
  const uint32_t b = (uint32_t)atoi(argv[2]);
  uint32_t L = 0;
  while ((1u << L) < b) ++L;

  or_dataset d;
  if (or_parse_csdb(buf, (size_t)n / 2, b, &d) == 0) {  /* truncated input must be rejected */
    fprintf(stderr, "truncated buffer accepted\n");
    return 1;
  }
  or_free(&d);
  if (or_parse_csdb(buf, (size_t)n, b, &d) != 0) return 1;
  const uint32_t nb = d.num_bricks;
  uint32_t B = 1;
  while (B * B * B < nb) ++B;
  uint32_t st;
  for (uint32_t i = 0; i < nb; ++i) {
    uint8_t* nib = (uint8_t*)malloc((d.bricks[i].nsym + 2) / 2 + 1);
    or_rans_decode(&d, i, nib, &st);
    free(nib);
  }
  uint8_t tmp[8];
  or_rans_decode(&d, nb, tmp, &st);  /* out of range: status, no access */
  const uint64_t per = pyr_off(L + 1);
  uint64_t* pyr = (uint64_t*)calloc(per * nb, sizeof(uint64_t));
  for (uint32_t i = 0; i < nb; ++i)
    for (uint32_t t = 0; t <= L; ++t) or_decode_dataset_brick(&d, i, t, pyr + per * i + pyr_off(t), &st);
  uint64_t dummy[1];
  or_decode_dataset_brick(&d, nb, 0, dummy, &st);
  uint64_t* pyr2 = (uint64_t*)calloc(per * nb, sizeof(uint64_t));
  or_decode_all_pyramid(&d, pyr2, 3, &st);
  if (memcmp(pyr, pyr2, per * nb * sizeof(uint64_t)) != 0) {
    fprintf(stderr, "threaded pyramid differs\n");
    return 1;
  }

  uint8_t* empty = (uint8_t*)calloc(nb, 1);
  uint8_t* lods = (uint8_t*)calloc(nb, 1);
  uint64_t* p0 = (uint64_t*)calloc(nb, sizeof(uint64_t));
  const uint64_t** labels = (const uint64_t**)calloc(nb, sizeof(uint64_t*));
  uint64_t* tfl = (uint64_t*)malloc(sizeof(uint64_t) * 64);
  uint32_t* tfc = (uint32_t*)malloc(sizeof(uint32_t) * 64);
  uint32_t ntf = 0;
  for (uint32_t i = 0; i < nb; ++i) {
    const or_brick* br = &d.bricks[i];
    int all0 = 1;
    for (uint32_t k = 0; k < br->P; ++k) all0 &= d.palette[br->pal_off + k] == 0;
    empty[i] = (uint8_t)all0;
    p0[i] = d.palette[br->pal_off];
    lods[i] = (uint8_t)((i * 7u) % (L + 1));
    labels[i] = (i % 5 == 4) ? NULL : pyr + per * i + pyr_off(lods[i]);  /* some bricks not resident */
    for (uint32_t k = 0; k < br->P && ntf < 64; ++k) {
      const uint64_t lab = d.palette[br->pal_off + k];
      int seen = lab == 0;
      for (uint32_t j = 0; j < ntf && !seen; ++j) seen = tfl[j] == lab;
      if (!seen) tfl[ntf++] = lab;
    }
  }
  for (uint32_t i = 1; i < ntf; ++i)  /* ascending */
    for (uint32_t j = i; j > 0 && tfl[j - 1] > tfl[j]; --j) {
      const uint64_t x = tfl[j];
      tfl[j] = tfl[j - 1];
      tfl[j - 1] = x;
    }
  for (uint32_t j = 0; j < ntf; ++j) tfc[j] = 0x60000000u | (j * 0x010305u);
  const float G = (float)(B * b);
  or_camera cams[2];
  memset(cams, 0, sizeof cams);
  cams[0].position[0] = G / 2; cams[0].position[1] = G / 2; cams[0].position[2] = -1.0f;
  cams[0].forward[2] = 1.0f; cams[0].right[0] = 1.0f; cams[0].up[1] = -1.0f; cams[0].ortho_half_height = G / 2;
  cams[1].position[0] = -0.5f * G; cams[1].position[1] = 1.3f * G; cams[1].position[2] = -0.7f * G;
  {
    float fwd[3] = {1.0f, -1.0f, 0.9f}, len = sqrtf(fwd[0] * fwd[0] + fwd[1] * fwd[1] + fwd[2] * fwd[2]);
    for (int k = 0; k < 3; ++k) cams[1].forward[k] = fwd[k] / len;
    cams[1].right[0] = 0.6f; cams[1].right[2] = -0.8f;
    cams[1].up[1] = 1.0f;
    cams[1].tan_half_fov = 0.577f;
  }
  const uint32_t w = 48, h = 32;
  uint64_t* lab = (uint64_t*)malloc(sizeof(uint64_t) * w * h);
  uint64_t* vox = (uint64_t*)malloc(sizeof(uint64_t) * w * h);
  float* tt = (float*)malloc(sizeof(float) * w * h);
  uint32_t* vis = (uint32_t*)calloc(nb, sizeof(uint32_t));
  uint32_t* rgba = (uint32_t*)malloc(sizeof(uint32_t) * w * h);
  const float light[3] = {0.42f, 0.6f, -0.9f};
  for (int c = 0; c < 2; ++c) {
    or_render(B, b, empty, lods, labels, p0, &cams[c], w, h, lab, vox, tt, vis);
    or_render_rgba(B, b, empty, lods, labels, p0, &cams[c], w, h, tfl, tfc, ntf, light, 0.5f, 0.8f, rgba);
  }
  for (uint32_t l = 0; l <= L; ++l) (void)or_brick_lod(1.0 * l, 2.0, 3.0, b, 0.0, 0.0, 0.0, 1.0, 1080.0, L);
  free(lab); free(vox); free(tt); free(vis); free(rgba);
  free(empty); free(lods); free(p0); free(labels); free(tfl); free(tfc);
  free(pyr); free(pyr2);
  or_free(&d);
  free(buf);
  printf("ok\n");
  return 0;
}
