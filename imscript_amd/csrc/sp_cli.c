/*
 * sp_cli.c -- the simpois, amle and heatd command lines and the image I/O
 * they need.
 *
 * main_simpois keeps the reference's flags, defaults, stdin/stdout
 * handling, mask semantics and exit codes (simpois.c:351-439), and calls
 * the GPU solver through poisson_solver_separable.  main_amle and amle do
 * the same for amle.c:203-246 and :325-366 through sp_amle_host, main_heatd
 * for heatd.c:109-150 through linear_diffusion_separable.
 *
 * Image I/O is the subset of iio that runs without external libraries:
 *   read : NPY 1.0 (iio.c:3216-3295), PFM (iio.c:2599-2616, see below)
 *   write: NPY '<f4' shape (h, w, pd) (iio.c:4337-4373), PFM for a ".pfm"
 *          suffix (iio.c:4151-4164); anything else, and pipes, get NPY
 *          (iio.c:6495); a terminal on stdout gets a one-line summary
 *          (iio.c:6171-6185, without the sixel preview).
 * Pixels are converted to planar float as iio_read_image_float_split does
 * (iio.c:1416-1420, 5763-5771).
 */
#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "../../include/simpois_hip.h"
#include "../../include/simpois_iio.h"
#include "sp_codecs.h"

/* ------------------------------------------------------------ failures */

static const char *g_cmdline = "simpois";

/* message and exit(-1), the behaviour of iio's fail() (iio.c:321-350) */
static void io_fail(const char *fmt, ...)
{
	va_list ap;
	fprintf(stderr, "\nIIO_ERROR(\"%s\"): ", g_cmdline);
	va_start(ap, fmt);
	vfprintf(stderr, fmt, ap);
	va_end(ap);
	fprintf(stderr, "\n\n");
	fflush(NULL);
	exit(-1);
}

static void *xalloc(size_t n)
{
	void *p = malloc(n ? n : 1);
	if (!p) {
		fprintf(stderr, "xmalloc: can not malloc %zu bytes\n", n);
		exit(1);
	}
	return p;
}

/* --------------------------------------------------------- option parse */

/* Remove "-o VALUE" from argv and return VALUE, or the default.  Same
 * contract as pick_option (pickopt.c:7-22): the scan starts at argv[0]
 * and the last argument cannot be an option name. */
static char *take_option(int *argc, char ***argv, const char *name, char *dflt)
{
	int n = *argc;
	char **v = *argv;
	int skip = dflt ? 1 : 0;
	for (int i = 0; i < n - skip; i++) {
		if (v[i][0] != '-' || strcmp(v[i] + 1, name)) continue;
		char *val = skip ? v[i + 1] : v[i] + 1;
		*argc -= skip + 1;
		for (int j = i; j < n - skip; j++)
			v[j] = v[j + skip + 1];
		return val;
	}
	return dflt;
}

/* ------------------------------------------------------------ reading */

typedef struct raw_image {
	int w, h, pd;
	float *px;      /* interleaved, w*h*pd */
} raw_image;

static FILE *open_in(const char *name)
{
	if (!strcmp(name, "-")) return stdin;
	FILE *f = fopen(name, "rb");
	if (!f) io_fail("can not open file \"%s\" in mode \"rb\"", name);
	return f;
}

static void read_exact(FILE *f, void *buf, size_t n)
{
	if (fread(buf, 1, n, f) != n) io_fail("input file ended before expected");
}

/* sample types (T_*) of NPY descr strings (iio.c:3266-3283) are those of
 * sp_codecs.h */

static int npy_type(const char *d, int *bytes, int *cplx)
{
	*cplx = 0;
	if (*d == '<' || *d == '>' || *d == '=' || *d == '|') d++;
	static const struct { const char *s; int t, b, c; } tab[] = {
		{"f8", T_F8, 8, 0}, {"f4", T_F4, 4, 0}, {"u1", T_U1, 1, 0},
		{"u2", T_U2, 2, 0}, {"u4", T_U4, 4, 0}, {"u8", T_U8, 8, 0},
		{"i1", T_I1, 1, 0}, {"i2", T_I2, 2, 0}, {"i4", T_I4, 4, 0},
		{"i8", T_I8, 8, 0}, {"c8", T_F4, 4, 1}, {"c16", T_F8, 8, 1},
		{"b1", T_I1, 1, 0}, {"B1", T_U1, 1, 0} };
	for (size_t k = 0; k < sizeof tab / sizeof *tab; k++)
		if (!strcmp(d, tab[k].s)) {
			*bytes = tab[k].b;
			*cplx = tab[k].c;
			return tab[k].t;
		}
	return -1;
}

static float sample_to_float(const unsigned char *p, int type)
{
	switch (type) {
	case T_F4: { float v; memcpy(&v, p, 4); return v; }
	case T_F8: { double v; memcpy(&v, p, 8); return (float)v; }
	case T_U1: return (float)p[0];
	case T_U2: { uint16_t v; memcpy(&v, p, 2); return (float)v; }
	case T_U4: { uint32_t v; memcpy(&v, p, 4); return (float)v; }
	case T_U8: { uint64_t v; memcpy(&v, p, 8); return (float)v; }
	case T_I1: return (float)(int8_t)p[0];
	case T_I2: { int16_t v; memcpy(&v, p, 2); return (float)v; }
	case T_I4: { int32_t v; memcpy(&v, p, 4); return (float)v; }
	default: { int64_t v; memcpy(&v, p, 8); return (float)v; }
	}
}

/* NPY 1.0 after the 4-byte magic "\x93NUM" (iio.c:3216-3295).  Like iio,
 * the byte-order character is skipped, not honoured. */
static int read_npy(FILE *f, raw_image *im, unsigned char **keep_raw,
		int *raw_type)
{
	unsigned char s[6];
	read_exact(f, s, 6);
	if (s[2] != 1 || s[3] != 0) io_fail("only NPY 1.0 is supported");
	int hl = s[4] + 256 * s[5];
	char *hdr = xalloc(hl + 1);
	read_exact(f, hdr, hl);
	hdr[hl] = 0;
	char descr[16] = {0}, order[16] = {0};
	int dims[3] = {0, 0, 0}, nd = 0;
	const char *d = strstr(hdr, "'descr'"), *o = strstr(hdr, "'fortran_order'"),
		*sh = strstr(hdr, "'shape'");
	if (!d || !o || !sh) io_fail("badly formed npy header");
	if (sscanf(d, "'descr': '%15[^']'", descr) != 1) io_fail("badly formed npy header");
	if (sscanf(o, "'fortran_order': %15[^,}]", order) != 1) io_fail("badly formed npy header");
	const char *par = strchr(sh, '(');
	if (!par) io_fail("badly formed npy header");
	par++;
	while (nd < 3) {
		while (*par == ' ') par++;
		if (*par == ')' || !isdigit((unsigned char)*par)) break;
		dims[nd++] = (int)strtol(par, (char **)&par, 10);
		while (*par == ' ') par++;
		if (*par == ',') par++;
	}
	free(hdr);
	if (nd < 1) io_fail("badly formed npy header");
	int h = dims[0], w = nd > 1 ? dims[1] : 1, pd = nd > 2 ? dims[2] : 1;
	if (nd < 2) { w = dims[0]; h = 1; }
	if (h == 1 && w > 1 && pd > 1) { h = w; w = pd; pd = 1; }   /* :3240-3245 */
	int fortran = order[0] == 'T';
	if (fortran) { int t = h; h = w; w = t; }
	int bytes, cplx, type = npy_type(descr, &bytes, &cplx);
	if (type < 0) io_fail("unrecognized npy type \"%s\"", descr);
	if (cplx) pd *= 2;
	size_t n = (size_t)w * h * pd;
	unsigned char *raw = xalloc(n * bytes);
	size_t got = fread(raw, bytes, n, f);
	if (got != n) {
		fprintf(stderr, "IIO WARNING: npy file smaller than expected\n");
		memset(raw + got * bytes, 0, (n - got) * bytes);
	}
	if (fortran) {   /* stored column-major: transpose to row-major */
		unsigned char *t = xalloc(n * bytes);
		for (int j = 0; j < h; j++)
		for (int i = 0; i < w; i++)
			memcpy(t + ((size_t)j * w + i) * pd * bytes,
				raw + ((size_t)i * h + j) * pd * bytes, (size_t)pd * bytes);
		free(raw);
		raw = t;
	}
	im->w = w; im->h = h; im->pd = pd;
	im->px = xalloc(n * sizeof(float));
	for (size_t k = 0; k < n; k++)
		im->px[k] = sample_to_float(raw + k * bytes, type);
	if (keep_raw) {   /* keep raw samples for the mask conversion */
		*raw_type = type;
		*keep_raw = raw;
	} else {
		free(raw);
	}
	return 0;
}

/* PFM: "Pf" (1 channel) or "PF" (3 channels), "w h\nscale\n", float rows.
 * iio.c:2599-2616 declares every PFM as 2-channel, which misreads both
 * kinds; this reader follows the header, as the matching PFM writer
 * (iio.c:4151-4164) intends. */
static void read_pfm(FILE *f, int upper, raw_image *im)
{
	int w, h;
	float scale;
	int c = fgetc(f);
	if (!isspace(c)) io_fail("bad PFM header");
	if (fscanf(f, "%d %d\n%g", &w, &h, &scale) != 3) io_fail("bad PFM header");
	c = fgetc(f);
	if (!isspace(c)) io_fail("bad PFM header");
	int pd = upper ? 3 : 1;
	size_t n = (size_t)w * h * pd;
	im->w = w; im->h = h; im->pd = pd;
	im->px = xalloc(n * sizeof(float));
	read_exact(f, im->px, n * sizeof(float));
}

/* the rest of a stream after the n bytes already read into head */
static unsigned char *slurp(FILE *f, const unsigned char *head, size_t n, size_t *len)
{
	size_t cap = 1 << 16, k = n;
	unsigned char *b = xalloc(cap);
	memcpy(b, head, n);
	for (;;) {
		if (k == cap) {
			unsigned char *t = realloc(b, cap *= 2);
			if (!t) io_fail("out of memory reading an image");
			b = t;
		}
		const size_t r = fread(b + k, 1, cap - k, f);
		k += r;
		if (!r) break;
	}
	*len = k;
	return b;
}

/* PNG or TIFF through sp_codecs.c, samples converted as
 * iio_convert_samples does (iio.c:980-1090: plain casts to float) */
static void read_coded(FILE *f, const unsigned char *head, int png, raw_image *im,
		unsigned char **keep_raw, int *raw_type)
{
	size_t n;
	unsigned char *b = slurp(f, head, 4, &n);
	codec_image c;
	char err[256];
	if ((png ? codec_read_png : codec_read_tiff)(b, n, &c, err, sizeof err))
		io_fail("%s", err);
	free(b);
	const size_t ns = (size_t)c.w * c.h * c.pd;
	const int sb = codec_type_bytes(c.type);
	im->w = c.w; im->h = c.h; im->pd = c.pd;
	im->px = xalloc(ns * sizeof(float));
	for (size_t k = 0; k < ns; k++)
		im->px[k] = sample_to_float(c.data + k * sb, c.type);
	if (keep_raw) {
		*raw_type = c.type;
		*keep_raw = c.data;
	} else {
		free(c.data);
	}
}

static void read_any(const char *name, raw_image *im, unsigned char **raw,
		int *raw_type)
{
	FILE *f = open_in(name);
	unsigned char m[4];
	read_exact(f, m, 2);
	if (raw) *raw = NULL;
	if (m[0] == 'P' && (m[1] == 'f' || m[1] == 'F')) {
		read_pfm(f, m[1] == 'F', im);
	} else {
		read_exact(f, m + 2, 2);
		if (m[0] == 0x93 && m[1] == 'N' && m[2] == 'U' && m[3] == 'M')
			read_npy(f, im, raw, raw_type);
		else if ((m[0] == 'I' && m[1] == 'I' && m[2] == 42 && m[3] == 0) ||
				(m[0] == 'M' && m[1] == 'M' && m[2] == 0 && m[3] == 42))
			read_coded(f, m, 0, im, raw, raw_type);
		else if (m[0] == 0x89 && m[1] == 'P' && m[2] == 'N' && m[3] == 'G')
			read_coded(f, m, 1, im, raw, raw_type);
		else
			io_fail("unsupported image format (NPY, PFM, TIFF and PNG are supported)");
	}
	if (f != stdin) fclose(f);
}

/* planar float image, iio_read_image_float_split (iio.c:5763-5771) */
static float *read_split(const char *name, int *w, int *h, int *pd)
{
	raw_image im;
	read_any(name, &im, NULL, NULL);
	size_t n = (size_t)im.w * im.h;
	float *out = xalloc(n * im.pd * sizeof(float));
	for (size_t i = 0; i < n; i++)
		for (int l = 0; l < im.pd; l++)
			out[n * l + i] = im.px[im.pd * i + l];
	free(im.px);
	*w = im.w; *h = im.h; *pd = im.pd;
	return out;
}

/* gray value of a colour sample triple, computed in double and stored in
 * the sample type, as iio_hacky_uncolorize does (iio.c:1230-1300) */
static float gray_of(const unsigned char *raw, int type, size_t k, int pd)
{
	int bytes = type == T_F8 || type == T_U8 || type == T_I8 ? 8
		: type == T_F4 || type == T_U4 || type == T_I4 ? 4
		: type == T_U2 || type == T_I2 ? 2 : 1;
	double a = sample_to_float(raw + (k * pd + 0) * bytes, type);
	double b = sample_to_float(raw + (k * pd + 1) * bytes, type);
	double c = sample_to_float(raw + (k * pd + 2) * bytes, type);
	double v = .299 * a + .587 * b + .114 * c;
	switch (type) {
	case T_U1: return (float)(uint8_t)v;
	case T_U2: return (float)(uint16_t)v;
	case T_F4: return (float)v;
	default: io_fail("uncolorize type not supported");
	}
	return 0;
}

/* scalar mask, iio_read_image_float (iio.c:6029-6048) */
static float *read_mask(const char *name, int *w, int *h)
{
	raw_image im;
	unsigned char *raw = NULL;
	int type = T_F4;
	read_any(name, &im, &raw, &type);
	size_t n = (size_t)im.w * im.h;
	float *out = xalloc(n * sizeof(float));
	if (im.pd == 1) {
		memcpy(out, im.px, n * sizeof(float));
	} else if (im.pd == 3 || im.pd == 4) {
		for (size_t i = 0; i < n; i++) {
			if (raw) out[i] = gray_of(raw, type, i, im.pd);
			else out[i] = (float)(.299 * im.px[im.pd * i]
					+ .587 * im.px[im.pd * i + 1]
					+ .114 * im.px[im.pd * i + 2]);
		}
	} else {
		io_fail("could not read image");
	}
	free(raw);
	free(im.px);
	*w = im.w; *h = im.h;
	return out;
}

/* ------------------------------------------------------------ writing */

static int has_suffix(const char *s, const char *suf)
{
	size_t a = strlen(s), b = strlen(suf);
	return a >= b && !strcmp(s + a - b, suf);
}

static FILE *open_out(const char *name)
{
	if (!strcmp(name, "-")) return stdout;
	FILE *f = fopen(name, "wb");
	if (!f) io_fail("can not open file \"%s\" in mode \"w\"", name);
	return f;
}

/* NPY header exactly as iio_write_image_as_npy lays it out
 * (iio.c:4337-4373): padded with spaces to a multiple of 16, ending "\n" */
static void write_npy(const char *name, const float *px, int w, int h, int pd)
{
	char buf[1000];
	buf[0] = (char)0x93;
	memcpy(buf + 1, "NUMPY", 5);
	buf[6] = 1; buf[7] = 0;
	int n = 10;
	n += snprintf(buf + n, sizeof buf - n,
		"{'descr': '<f4', 'fortran_order': False, 'shape': (%d, %d, %d)}",
		h, w, pd);
	int m = ((n + 15 + 1) / 16) * 16;
	buf[8] = (char)((m - 10) % 256);
	buf[9] = (char)((m - 10) / 256);
	for (int i = n; i < m - 1; i++) buf[i] = ' ';
	buf[m - 1] = '\n';
	FILE *f = open_out(name);
	fwrite(buf, 1, m, f);
	fwrite(px, sizeof(float), (size_t)w * h * pd, f);
	if (f != stdout) fclose(f); else fflush(f);
}

static void write_pfm(const char *name, const float *px, int w, int h, int pd)
{
	FILE *f = open_out(name);
	fprintf(f, "P%c\n%d %d\n%g\n", pd > 1 ? 'F' : 'f', w, h, -1.0f);
	fwrite(px, sizeof(float), (size_t)w * h * pd, f);
	if (f != stdout) fclose(f); else fflush(f);
}

/* float to 8 / 16 bits as convert_datum does (iio.c:1036-1039): 0.5 added
 * in double, clamped, truncated; NaN becomes 0 */
static unsigned char to_u8(float v)
{
	const double d = 0.5 + v;
	return d > 0 ? (d < 0xff ? (unsigned char)d : 0xff) : 0;
}

static unsigned short to_u16(float v)
{
	const double d = 0.5 + v;
	return d > 0 ? (d < 0xffff ? (unsigned short)d : 0xffff) : 0;
}

static void write_png(const char *name, const float *px, int w, int h, int pd, int b16)
{
	if (pd < 1 || pd > 4) io_fail("can not save %d-dimensional samples as PNG", pd);
	const size_t ns = (size_t)w * h * pd;
	unsigned char *p8 = b16 ? NULL : xalloc(ns);
	unsigned short *p16 = b16 ? xalloc(ns * 2) : NULL;
	for (size_t k = 0; k < ns; k++) {
		if (b16) p16[k] = to_u16(px[k]);
		else p8[k] = to_u8(px[k]);
	}
	FILE *f = open_out(name);
	if (codec_write_png(f, p8, p16, w, h, pd)) io_fail("png write error");
	if (f != stdout) fclose(f); else fflush(f);
	free(p8);
	free(p16);
}

/* float TIFF, LZW below 2000 x 2000 pixels unless IIOTIFF_PLAIN is set
 * (iio.c:4043-4047) */
static void write_tiff(const char *name, const float *px, int w, int h, int pd)
{
	const int lzw = (long)w * h < 2000L * 2000 && !getenv("IIOTIFF_PLAIN");
	FILE *f = open_out(name);
	if (codec_write_tiff(f, px, w, h, pd, lzw)) io_fail("error writing TIFF file");
	if (f != stdout) fclose(f); else fflush(f);
}

/* the writer dispatch of iio_write_image_default (iio.c:6166-6497) for the
 * formats kept here, on interleaved floats */
static void write_vec(const char *name, const float *px, int w, int h, int pd)
{
	if (!strcmp(name, "-") && isatty(1)) {
		printf("IMAGE %dx%d,%d %s\n", w, h, pd, "FLOAT");
		return;
	}
	if (has_suffix(name, ".pfm") && (pd == 1 || pd == 3))
		write_pfm(name, px, w, h, pd);
	else if (has_suffix(name, ".npy"))
		write_npy(name, px, w, h, pd);
	else if (has_suffix(name, ".tiff") || has_suffix(name, ".tif") ||
			has_suffix(name, ".TIFF") || has_suffix(name, ".TIF"))
		write_tiff(name, px, w, h, pd);
	else if (!strncmp(name, "TIFF:", 5))
		write_tiff(name + 5, px, w, h, pd);
	else if (!strncmp(name, "PNG:", 4))
		write_png(name + 4, px, w, h, pd, 0);
	else if (!strncmp(name, "PNG16:", 6))
		write_png(name + 6, px, w, h, pd, 1);
	else if (has_suffix(name, ".png") || has_suffix(name, ".PNG"))
		write_png(name, px, w, h, pd, 0);
	else
		write_npy(name, px, w, h, pd);
}

/* iio_write_image_float_split (iio.c:6525-6531): planar to interleaved */
static void write_split(const char *name, const float *planar, int w, int h, int pd)
{
	size_t n = (size_t)w * h;
	float *px = xalloc(n * pd * sizeof(float));
	for (int l = 0; l < pd; l++)
		for (size_t i = 0; i < n; i++)
			px[pd * i + l] = planar[n * l + i];
	write_vec(name, px, w, h, pd);
	free(px);
}

/* --------------------------------------------------------------- help */

/* The help texts are the reference's (simpois.c:307-347), kept verbatim
 * because they are part of the command-line interface. */
static const char *help_usage = "usage:\n\t"
"simpois [options] [-m mask.png] [-f data.npy] [-i in.npy] [-o out.npy]";
static const char *help_oneliner = "a simple Poisson solver for images";
static const char *help_version = "simpois 1.0\n\nWritten by mnhrdt";
static const char *help_long =
"Simpois is a simple Poisson solver for image processing.\n"
"\n"
"This program can solve Poisson, Laplace and Biharmonic equations\n"
"over an arbitrary part of the image domain, with arbitrary data terms\n"
"and boundary conditions.  Fast approximations to the solution can be\n"
"obtained by setting the parameters.  Solving over the whole image domain\n"
"is discouraged, as it is always better to use fft-based methods.\n"
"\n"
"Usage: simpois [options] -i in.npy -o out.npy\n"
"   or: simpois [options] -i in.npy > out.npy\n"
"   or: cat in.npy | simpois [options] -o out.npy\n"
"   or: cat in.npy | simpois [options] > out.npy\n"
"\n"
"Options:\n"
" -i IMG\tInput image with boundary data (default=stdin)\n"
" -f IMG\tPoisson data term (default=zeros)\n"
" -m IMG\tMask for the region of interest (default=zeros)\n"
" -o IMG\tOutput image (default=stdout)\n"
" -t 0.25\tTime step for Gauss-Seidel iterations (negative=biharmonic)\n"
" -n 10\tNumber of Gauss-Seidel iterations\n"
" -s 99\tMaximum number of multi-scale octaves\n"
" -c 0\tNumber of Conjugate Gradient iterations\n"
" -h\tdisplay short help message\n"
" --help\tdisplay longer help message\n"
"\n"
"Examples:\n"
" cat in.npy | simpois > out.npy           Fill NANs by Laplace equation\n"
" cat in.npy | simpois -n 1 > out.npy      Fill NANs, fast (one iteration)\n"
" cat in.npy | simpois -t -0.08 > out.npy  Fill NANs, smooth (Biharmonic)\n"
" simpois -i in.npy -o out.npy             Laplace, with explicit data\n"
" simpois -m mask.png ...                  Use mask instead of NANs\n"
" simpois -f lap.npy ...                   Poisson editor\n"
"\n"
"Report bugs to <enric.meinhardt@ens-paris-saclay.fr>.";

/* help flags, help_stuff.c:23-34 (the --man* pages need help2man and are
 * not provided) */
static void maybe_help(const char *s)
{
	if (!s || !strcmp(s, "-h")) { puts(help_usage); exit(0); }
	if (!strcmp(s, "-?")) { puts(help_oneliner); exit(0); }
	if (!strcmp(s, "--help")) { puts(help_long); exit(0); }
	if (!strcmp(s, "--version")) { puts(help_version); exit(0); }
	if (!strncmp(s, "--man", 5)) {
		fprintf(stderr, "simpois: man pages are not available\n");
		exit(1);
	}
	if (!strcmp(s, "--help-oneliner")) puts(help_oneliner);
}

/* --------------------------------------------------------------- main */

int main_simpois(int argc, char *argv[])
{
	static char cmd[1024];
	cmd[0] = 0;
	for (int i = 0; i < argc; i++) {
		if (i) strncat(cmd, " ", sizeof cmd - strlen(cmd) - 1);
		strncat(cmd, argv[i], sizeof cmd - strlen(cmd) - 1);
	}
	g_cmdline = cmd;
	if (argc == 2) maybe_help(argv[1]);

	float tstep = atof(take_option(&argc, &argv, "t", "0.25"));
	float niter = atof(take_option(&argc, &argv, "n", "10"));
	float nscal = atof(take_option(&argc, &argv, "s", "99"));
	float cgrad = atof(take_option(&argc, &argv, "c", "0"));
	char *fi = take_option(&argc, &argv, "i", "-");
	char *fo = take_option(&argc, &argv, "o", "-");
	char *fm = take_option(&argc, &argv, "m", "");
	char *ff = take_option(&argc, &argv, "f", "");
	if (!strcmp(ff, "-") && !strcmp(fi, "-"))
		fi = "NAN";

	if (argc > 1) {   /* simpois.c:369-386 */
		fprintf(stderr, "Usage:\n\t%s [options]\n", *argv);
		fprintf(stderr, "\nComputes approximate solutions of Poisson"
			" or Laplace equations for images\n");
		fprintf(stderr, "\nOptions with their default values:\n"
			"\t-i stdin   Input image with boundary data\n"
			"\t-f (zeros) Optional image with Poisson data term\n"
			"\t-m (zeros) Optional image with region of interest\n"
			"\t-o stdout  Output image\n"
			"\t-t 0.25    Time step for Gauss-Seidel iterations\n"
			"\t-n 10      Number of Gauss-Seidel iterations\n"
			"\t-s 99      Number of Multi-Scale octaves\n"
			"\t-c 0       Number of Conjugate Gradient iterations\n"
		);
		fprintf(stderr, "\nNote: NAN values in the input image"
				" are added to the region of interest\n");
		return 1;
	}

	int w = 0, h = 0, pd = 0;
	float *img_i = NULL;
	if (*fi && strcmp(fi, "NAN"))
		img_i = read_split(fi, &w, &h, &pd);

	float *img_f = NULL;
	if (*ff) {
		int w2, h2, pd2;
		img_f = read_split(ff, &w2, &h2, &pd2);
		if (!img_i) {
			w = w2; h = h2; pd = pd2;
			img_i = xalloc((size_t)w * h * pd * sizeof(float));
			for (size_t k = 0; k < (size_t)w * h * pd; k++) img_i[k] = NAN;
		} else if (w != w2 || h != h2 || pd != pd2) {
			return fprintf(stderr, "input sizes mismatch (i,f)");
		}
	}
	if (!img_i) io_fail("could not read image");

	float *img_m = NULL;
	if (*fm) {
		int w2, h2;
		img_m = read_mask(fm, &w2, &h2);
		if (w != w2 || h != h2)
			return fprintf(stderr, "input sizes mismatch (i,m)");
	}

	float *out = xalloc((size_t)w * h * pd * sizeof(float));
	if (img_m)   /* mask > 0 makes a hole in every channel, :421-424 */
		for (size_t k = 0; k < (size_t)w * h * pd; k++)
			if (img_m[k % ((size_t)w * h)] > 0)
				img_i[k] = NAN;

	poisson_solver_separable(out, img_i, img_f, w, h, pd,
			tstep, niter, nscal, cgrad);

	write_split(fo, out, w, h, pd);
	free(out);
	free(img_i);
	free(img_m);
	free(img_f);
	return 0;
}

/* ================================================================ amle */

/* SMART_PARAMETER (smapa.h:19-42): getenv + sscanf("%lf") on first use,
 * with the same message on stderr */
static double amle_knob(const char *name, double dflt, int *known, double *val)
{
	if (*known) return *val;
	fprintf(stderr, "scanning the environment for \"%s\"... ", name);
	const char *sv = getenv(name);
	double y;
	if (sv && sscanf(sv, "%lf", &y) == 1) {
		fprintf(stderr, "got value %g\n", y);
		*val = y;
	} else {
		fprintf(stderr, "kept default value %g\n", dflt);
		*val = dflt;
	}
	*known = 1;
	return *val;
}

static int g_niter_known, g_nn_known;
static double g_niter_val, g_nn_val;

/* fail() of fail.c:62-80: message and exit(-1) */
static void amle_fail(const char *msg)
{
	fprintf(stderr, "\nFAIL(\"%s\"): %s\n\n", g_cmdline, msg);
	fflush(NULL);
	exit(-1);
}

/* amle(y, x, w, h) of amle.c:203-246 on the GPU: AMLE_NITER and AMLE_NN
 * (smart parameters, defaults 100 and 4), AMLE_INIT (an image file whose
 * values start tsup at the holes), the diagnostic line every tenth sweep
 * on stderr (SIMPOIS_HIP_QUIET=1 leaves it out and skips the tinf sweeps
 * it needs).  AMLE_NN outside 1..32 is refused (the reference reads past
 * its neighbour table there). */
void amle(float *y, float *x, int w, int h)
{
	const size_t n = (size_t)w * h;
	int holes = 0;
	for (size_t k = 0; k < n && !holes; k++) holes = isnan(x[k]);
	const double vit = amle_knob("AMLE_NITER", 100, &g_niter_known, &g_niter_val);
	const int niter = vit > 0 ? (int)ceil(vit) : 0;
	int nn = 4;
	if (niter > 0 && holes)
		nn = (int)amle_knob("AMLE_NN", 4, &g_nn_known, &g_nn_val);
	float *init = NULL;
	const char *fi = getenv("AMLE_INIT");
	if (fi) {
		int ww, hh;
		init = read_mask(fi, &ww, &hh);
		if ((long)ww * hh != (long)w * h) amle_fail("init size mismatch");
	}
	const char *q = getenv("SIMPOIS_HIP_QUIET");
	const int trace = !(q && atoi(q));
	if (sp_amle_host(y, x, init, w, h, 1, nn, niter, trace)) {
		fprintf(stderr, "amle: %s\n", sp_last_error());
		exit(1);
	}
	free(init);
}

/* help texts of amle.c:292-323, kept verbatim as part of the interface */
static const char *amle_help_usage = "usage:\n\t"
"simpois [in [mask [out]]]";
static const char *amle_help_oneliner = "Absolutely Minimizing Lipschitz Extension";
static const char *amle_help_version = "amle 1.0\n\nWritten by mnhrdt";
static const char *amle_help_long =
"Interpolate by AMLE (Absolutely Minimizing Lipschitz Extension).\n"
"\n"
"The pixels to fill-in are specified by NAN values in the input image\n"
"or by a separate user-provided mask.  The interpolated is defined\n"
"as that whose maximal slope is the lowest possible, recursively in all\n"
"subsets on the region of interest.\n"
"\n"
"Usage: amle in.npy mask.png out.npy\n"
"   or: amle in.npy mask.png > out.npy\n"
"   or: amle in.npy > out.npy\n"
"   or: cat in.npy | amle > out.npy\n"
"\n"
"Options:\n"
" -h\tdisplay short help message\n"
" --help\tdisplay longer help message\n"
"\n"
"Environement:\n"
" AMLE_NN\tNeigborhood size (4, 6, 8, 16, 26 or 32, default=4)\n"
" AMLE_TAU\tTimestep for the iterations (default=0.25)\n"
" AMLE_NITER\tNumber of iterations at each scale (default=100)\n"
" AMLE_INIT\tInitial image for the iterations (default=zeros)\n"
"\n"
"Examples:\n"
" amle in.npy > out.npy               Fill-in the holes in a DSM\n"
" AMLE_NN=32 amle in.npy > out.npy    More precise, slower solver\n"
"\n"
"Report bugs to <enric.meinhardt@ens-paris-saclay.fr>.";

static void amle_maybe_help(const char *s)
{
	if (!s || !strcmp(s, "-h")) { puts(amle_help_usage); exit(0); }
	if (!strcmp(s, "-?")) { puts(amle_help_oneliner); exit(0); }
	if (!strcmp(s, "--help")) { puts(amle_help_long); exit(0); }
	if (!strcmp(s, "--version")) { puts(amle_help_version); exit(0); }
	if (!strncmp(s, "--man", 5)) {
		fprintf(stderr, "amle: man pages are not available\n");
		exit(1);
	}
	if (!strcmp(s, "--help-oneliner")) puts(amle_help_oneliner);
}

/* The amle command line (amle.c:325-366): amle [in [mask [out]]], stdin
 * and stdout by default, mask > 0 makes a hole, scalar images (colour
 * inputs are converted to gray as iio_read_image_float does). */
int main_amle(int c, char *v[])
{
	static char cmd[1024];
	cmd[0] = 0;
	for (int i = 0; i < c; i++) {
		if (i) strncat(cmd, " ", sizeof cmd - strlen(cmd) - 1);
		strncat(cmd, v[i], sizeof cmd - strlen(cmd) - 1);
	}
	g_cmdline = cmd;
	if (c == 2) amle_maybe_help(v[1]);
	if (c > 4 || (c > 1 && v[1][0] == '-' && v[1][1] == 'h')) {
		fprintf(stderr, "usage:\n\t%s [in [mask [out]]]\n", *v);
		return 1;
	}
	const char *fin = c > 1 ? v[1] : "-";
	const char *fmask = c > 2 ? v[2] : "";
	const char *fout = c > 3 ? v[3] : "-";
	int w[2], h[2];
	float *in = read_mask(fin, w, h);
	float *mask = NULL;
	if (fmask[0]) {
		mask = read_mask(fmask, w + 1, h + 1);
		if (w[0] != w[1] || h[0] != h[1])
			amle_fail("image and mask file size mismatch");
	}
	const size_t n = (size_t)w[0] * h[0];
	float *out = xalloc(n * sizeof(float));
	if (mask)
		for (size_t k = 0; k < n; k++)
			if (mask[k] > 0) in[k] = NAN;
	amle(out, in, w[0], h[0]);
	write_split(fout, out, w[0], h[0], 1);
	free(out);
	free(mask);
	free(in);
	return 0;
}

/* =============================================================== heatd */

/* The heatd command line (heatd.c:109-150): heatd [-n 1] [-t 0.25]
 * [-A field] [in [out]]; -g, -m and -D are accepted and ignored as there.
 * The field is an interleaved w x h x 3 image (a, b, c).  -n 0 writes the
 * input unchanged (the reference writes an unset buffer). */
int main_heatd(int c, char *v[])
{
	static char cmd[1024];
	cmd[0] = 0;
	for (int i = 0; i < c; i++) {
		if (i) strncat(cmd, " ", sizeof cmd - strlen(cmd) - 1);
		strncat(cmd, v[i], sizeof cmd - strlen(cmd) - 1);
	}
	g_cmdline = cmd;
	int nsteps = atoi(take_option(&c, &v, "n", "1"));
	float tstep = atof(take_option(&c, &v, "t", "0.25"));
	(void)take_option(&c, &v, "g", "");
	char *fa = take_option(&c, &v, "A", "");
	(void)take_option(&c, &v, "m", "");
	(void)take_option(&c, &v, "D", "");
	if (c != 1 && c != 2 && c != 3)
		return fprintf(stderr, "usage:\n\t%s [in [out]]\n", *v);
	const char *fin = c > 1 ? v[1] : "-";
	const char *fout = c > 2 ? v[2] : "-";
	int w, h, pd;
	float *x = read_split(fin, &w, &h, &pd);
	float *A = NULL;
	if (*fa) {
		raw_image im;
		read_any(fa, &im, NULL, NULL);
		if (im.w != w || im.h != h || im.pd != 3)
			return fprintf(stderr, "ERROR: bad A dimensions"
				"(%d %d %d) ((%d %d))\n", im.w, im.h, im.pd, w, h);
		A = im.px;
	}
	const size_t n = (size_t)w * h * pd;
	float *y = xalloc(n * sizeof(float));
	if (nsteps > 0) linear_diffusion_separable(y, x, w, h, pd, A, nsteps, tstep);
	else memcpy(y, x, n * sizeof(float));
	write_split(fout, y, w, h, pd);
	free(y);
	free(A);
	free(x);
	return 0;
}

/* ===================================================== iio entry points */

/* iio.c:5747-5760 */
float *iio_read_image_float_vec(const char *fname, int *w, int *h, int *pd)
{
	raw_image im;
	read_any(fname, &im, NULL, NULL);
	*w = im.w; *h = im.h; *pd = im.pd;
	return im.px;
}

/* iio.c:5763-5771 */
float *iio_read_image_float_split(const char *fname, int *w, int *h, int *pd)
{
	return read_split(fname, w, h, pd);
}

/* iio.c:6029-6048 */
float *iio_read_image_float(const char *fname, int *w, int *h)
{
	return read_mask(fname, w, h);
}

/* iio.c:6516-6523 */
void iio_write_image_float_vec(char *filename, float *x, int w, int h, int pd)
{
	write_vec(filename, x, w, h, pd);
}

/* iio.c:6525-6531 */
void iio_write_image_float_split(char *filename, float *x, int w, int h, int pd)
{
	write_split(filename, x, w, h, pd);
}

/* iio.c:6570-6576 */
void iio_write_image_float(char *filename, float *x, int w, int h)
{
	write_vec(filename, x, w, h, 1);
}
