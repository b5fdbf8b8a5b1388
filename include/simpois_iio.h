/*
 * simpois_iio.h -- the image I/O entry points of iio that the command
 * lines of this library use, with iio's signatures and behaviour, so that
 * a program written against iio.h can take its image I/O from here.
 *
 * Formats read (by content, from a file name or "-" for stdin): NPY 1.0,
 * PFM, TIFF (strips or tiles, chunky or planar; uncompressed, LZW,
 * Deflate or PackBits; predictors 1-3; 1-64-bit integer and 32/64-bit
 * float samples) and PNG (all colour types and depths, interlaced or not;
 * palettes and tRNS expanded as libpng's EXPAND transform does).
 * Formats written, by file name: ".pfm" (1 or 3 channels), ".npy", ".tif"
 * / ".tiff" / "TIFF:" prefix (float32, LZW below 2000x2000 pixels unless
 * IIOTIFF_PLAIN is set), ".png" / "PNG:" (8 bits) and "PNG16:" (16 bits)
 * with samples rounded and clamped as iio converts them; anything else and
 * pipes get NPY.  "-" writes to stdout (a terminal gets a one-line
 * summary).  A read or write failure prints a message and exits, as iio's
 * fail() does.
 */
#ifndef SIMPOIS_IIO_H
#define SIMPOIS_IIO_H

#ifdef __cplusplus
extern "C" {
#endif

/* interleaved samples, w*h*pd floats (malloc'd).  Replaces iio.c:5747-5760
 * (iio_read_image_float_vec), declared at iio.h:36. */
float *iio_read_image_float_vec(const char *fname, int *w, int *h, int *pd);

/* planar samples, plane l at +w*h*l.  Replaces iio.c:5763-5771
 * (iio_read_image_float_split), iio.h:42. */
float *iio_read_image_float_split(const char *fname, int *w, int *h, int *pd);

/* one channel: colour images are converted to gray (0.299, 0.587, 0.114 in
 * the sample type).  Replaces iio.c:6029-6048 (iio_read_image_float),
 * iio.h:33. */
float *iio_read_image_float(const char *fname, int *w, int *h);

/* Replace iio.c:6516-6523, 6525-6531 and 6570-6576
 * (iio_write_image_float_vec / _split / iio_write_image_float), iio.h:181-182. */
void iio_write_image_float_vec(char *filename, float *x, int w, int h, int pd);
void iio_write_image_float_split(char *filename, float *x, int w, int h, int pd);
void iio_write_image_float(char *filename, float *x, int w, int h);

#ifdef __cplusplus
}
#endif

#endif /* SIMPOIS_IIO_H */
