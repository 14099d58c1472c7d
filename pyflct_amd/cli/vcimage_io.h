/* vcimage file I/O for the flct / warp command-line programs.
 * Format: big-endian int32 magic 2136967593, int32 nx_idl, int32 ny_idl, then
 * float32 arrays with IDL x fastest (flct.c:1158-1268, flct.c:2016-2109).
 * Arrays handed to and from these functions are C-space doubles: nx = IDL y
 * size (rows), ny = IDL x size (columns). */
#ifndef FLCT_VCIMAGE_IO_H
#define FLCT_VCIMAGE_IO_H

#define VCIMAGE_MAGIC 2136967593

/* Reads `count` arrays; returns 0 on success, prints a message and returns
 * nonzero otherwise.  *arrs receives `count` malloc'd arrays. */
int vc_read(const char *who, const char *path, int count, int *nx, int *ny, double **arrs);

/* Writes `count` arrays narrowed to float32; returns 0 on success. */
int vc_write(const char *who, const char *path, int count, int nx, int ny,
             const double *const *arrs);

#endif
