#include "vcimage_io.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint32_t be32(const unsigned char *b) {
  return ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | (uint32_t)b[3];
}

static void put_be32(unsigned char *b, uint32_t v) {
  b[0] = (unsigned char)(v >> 24);
  b[1] = (unsigned char)(v >> 16);
  b[2] = (unsigned char)(v >> 8);
  b[3] = (unsigned char)v;
}

int vc_read(const char *who, const char *path, int count, int *nx, int *ny, double **arrs) {
  FILE *fp = fopen(path, "rb");
  unsigned char head[12];
  long long n, a, k;
  unsigned char *buf;
  if (!fp) {
    printf("%s: cannot open file %s\n", who, path);
    return 1;
  }
  if (fread(head, 1, 12, fp) != 12 || be32(head) != (uint32_t)VCIMAGE_MAGIC) {
    printf("%s: input file %s is not a vel_ccor i/o file\n", who, path);
    fclose(fp);
    return 1;
  }
  /* header holds IDL (nx, ny); C-space swaps them (flct.c:1192-1196) */
  *ny = (int)be32(head + 4);
  *nx = (int)be32(head + 8);
  if (*nx <= 0 || *ny <= 0) {
    printf("%s: bad dimensions in %s\n", who, path);
    fclose(fp);
    return 1;
  }
  n = (long long)(*nx) * (*ny);
  buf = (unsigned char *)malloc((size_t)n * 4);
  if (!buf) {
    printf("%s: memory request failed\n", who);
    fclose(fp);
    return 1;
  }
  for (a = 0; a < count; a++) {
    arrs[a] = (double *)malloc(sizeof(double) * (size_t)n);
    if (!arrs[a] || fread(buf, 4, (size_t)n, fp) != (size_t)n) {
      printf("%s: %s holds fewer than %d arrays\n", who, path, count);
      free(buf);
      fclose(fp);
      return 1;
    }
    for (k = 0; k < n; k++) {
      uint32_t u = be32(buf + 4 * k);
      float f;
      memcpy(&f, &u, 4);
      arrs[a][k] = (double)f;
    }
  }
  free(buf);
  fclose(fp);
  return 0;
}

int vc_write(const char *who, const char *path, int count, int nx, int ny,
             const double *const *arrs) {
  FILE *fp = fopen(path, "wb");
  unsigned char head[12];
  long long n = (long long)nx * ny, a, k;
  unsigned char *buf;
  if (!fp) {
    printf("%s: cannot open file %s\n", who, path);
    return 1;
  }
  put_be32(head, (uint32_t)VCIMAGE_MAGIC);
  put_be32(head + 4, (uint32_t)ny);
  put_be32(head + 8, (uint32_t)nx);
  buf = (unsigned char *)malloc((size_t)n * 4);
  if (!buf || fwrite(head, 1, 12, fp) != 12) {
    printf("%s: cannot write %s\n", who, path);
    free(buf);
    fclose(fp);
    return 1;
  }
  for (a = 0; a < count; a++) {
    for (k = 0; k < n; k++) {
      float f = (float)arrs[a][k];
      uint32_t u;
      memcpy(&u, &f, 4);
      put_be32(buf + 4 * k, u);
    }
    if (fwrite(buf, 4, (size_t)n, fp) != (size_t)n) {
      printf("%s: cannot write %s\n", who, path);
      free(buf);
      fclose(fp);
      return 1;
    }
  }
  free(buf);
  fclose(fp);
  return 0;
}
