/*
 * sp_codecs.h -- PNG and TIFF coding for the command lines' image I/O
 * (the part of iio that the reference takes from libpng and libtiff).
 * Internal to the library; not part of the public C ABI.
 */
#ifndef SP_CODECS_H
#define SP_CODECS_H

#include <stddef.h>
#include <stdio.h>

/* sample types (the NPY reader's too) */
enum { T_F4, T_F8, T_U1, T_U2, T_U4, T_U8, T_I1, T_I2, T_I4, T_I8 };

int codec_type_bytes(int type);

/* a decoded image: w*h*pd interleaved samples of one type, native order */
typedef struct codec_image {
	int w, h, pd, type;
	unsigned char *data;
} codec_image;

/* Decoders of a whole file in memory.  Return 0, or -1 with a message in
 * err.  PNG: 8- or 16-bit samples after libpng's PACKING | EXPAND
 * transforms (iio.c:1500-1556).  TIFF: the first image, samples as stored
 * (iio.c:1668-1915). */
int codec_read_png(const unsigned char *buf, size_t n, codec_image *im,
		char *err, size_t errn);
int codec_read_tiff(const unsigned char *buf, size_t n, codec_image *im,
		char *err, size_t errn);

/* PNG of w*h*pd interleaved samples (pd 1..4), 8-bit (px8) or 16-bit
 * (px16, native order); returns 0 or -1 */
int codec_write_png(FILE *f, const unsigned char *px8, const unsigned short *px16,
		int w, int h, int pd);

/* TIFF of w*h*pd interleaved float samples, chunky, LZW or uncompressed
 * (iio.c:4011-4081) */
int codec_write_tiff(FILE *f, const float *px, int w, int h, int pd, int lzw);

/* LZW (TIFF flavour: MSB-first codes, 9 to 12 bits, early change) of n
 * bytes; *out is allocated, returns its size or -1.  Exposed for tests. */
long codec_lzw_encode(const unsigned char *in, size_t n, unsigned char **out);
long codec_lzw_decode(const unsigned char *in, size_t n, unsigned char *out,
		size_t cap);

#endif
