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
 *   read : NPY 1.0 (iio.c:3216-3295), PFM (iio.c:2599-2616, see below),
 *          TIFF and PNG (sp_codecs.c); the names "zero:WxH[,pd]",
 *          "one:WxH[,pd]", "constant:V:WxH" (iio.c:5603-5634),
 *          "TRANS[list]:name" and the IIO_TRANS variable (flips, axis
 *          permutations, crops, pipes; iio.c:3800-3884, 5677-5704)
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

/* An image as read, before the conversion to float: the transformations
 * of IIO_TRANS and "TRANS[...]:" names and the mask's colour-to-gray step
 * act on the samples in their file type, as iio's do. */
typedef struct raw_image {
	int w, h, pd;
	int type;              /* sample type, T_* of sp_codecs.h */
	unsigned char *data;   /* interleaved samples, w*h*pd of that type */
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
static void read_npy(FILE *f, raw_image *im)
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
	im->type = type;
	im->data = raw;
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
	im->type = T_F4;
	im->data = xalloc(n * sizeof(float));
	read_exact(f, im->data, n * sizeof(float));
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

/* PNG or TIFF through sp_codecs.c, samples kept in their file type */
static void read_coded(FILE *f, const unsigned char *head, int png, raw_image *im)
{
	size_t n;
	unsigned char *b = slurp(f, head, 4, &n);
	codec_image c;
	char err[256];
	if ((png ? codec_read_png : codec_read_tiff)(b, n, &c, err, sizeof err))
		io_fail("%s", err);
	free(b);
	im->w = c.w; im->h = c.h; im->pd = c.pd;
	im->type = c.type;
	im->data = c.data;
}

/* one file or stream, format by its first bytes (read_image_f) */
static void read_core(const char *name, raw_image *im)
{
	FILE *f = open_in(name);
	unsigned char m[4];
	read_exact(f, m, 2);
	if (m[0] == 'P' && (m[1] == 'f' || m[1] == 'F')) {
		read_pfm(f, m[1] == 'F', im);
	} else {
		read_exact(f, m + 2, 2);
		if (m[0] == 0x93 && m[1] == 'N' && m[2] == 'U' && m[3] == 'M')
			read_npy(f, im);
		else if ((m[0] == 'I' && m[1] == 'I' && m[2] == 42 && m[3] == 0) ||
				(m[0] == 'M' && m[1] == 'M' && m[2] == 0 && m[3] == 42))
			read_coded(f, m, 0, im);
		else if (m[0] == 0x89 && m[1] == 'P' && m[2] == 'N' && m[3] == 'G')
			read_coded(f, m, 1, im);
		else
			io_fail("unsupported image format (NPY, PFM, TIFF and PNG are supported)");
	}
	if (f != stdin) fclose(f);
}

/* ------------------------------------ transformations of a read image */

/* Geometric transformations of an image as read (iio.c:757-960): they move
 * whole pixels of pd samples, in the samples' file type. */

static size_t px_bytes(const raw_image *im)
{
	return (size_t)im->pd * codec_type_bytes(im->type);
}

static void swap_px(raw_image *im, int i, int j, int a, int b)
{
	const size_t ps = px_bytes(im);
	unsigned char *p = im->data + ((size_t)j * im->w + i) * ps;
	unsigned char *q = im->data + ((size_t)b * im->w + a) * ps;
	for (size_t k = 0; k < ps; k++) {
		unsigned char t = p[k];
		p[k] = q[k];
		q[k] = t;
	}
}

static void flip_leftright(raw_image *im)
{
	for (int j = 0; j < im->h; j++)
		for (int i = 0; i < im->w / 2; i++)
			swap_px(im, i, j, im->w - i - 1, j);
}

static void flip_topdown(raw_image *im)
{
	for (int j = 0; j < im->h / 2; j++)
		for (int i = 0; i < im->w; i++)
			swap_px(im, i, j, i, im->h - j - 1);
}

static void transpose_px(raw_image *im)
{
	const size_t ps = px_bytes(im);
	const int w = im->w, h = im->h;
	unsigned char *t = xalloc((size_t)w * h * ps);
	for (int j = 0; j < h; j++)        /* new pixel (j, i) = old (i, j) */
		for (int i = 0; i < w; i++)
			memcpy(t + ((size_t)i * h + j) * ps,
				im->data + ((size_t)j * w + i) * ps, ps);
	free(im->data);
	im->data = t;
	im->w = h;
	im->h = w;
}

/* two orientation letters (iio.c:811-821): an upper-case first letter flips
 * left-right, an upper-case second one top-down, and the pair is transposed
 * when the first letter comes after the second in the alphabet */
static void reorient(raw_image *im, int ox, int oy)
{
	if (isupper(ox)) flip_leftright(im);
	if (isupper(oy)) flip_topdown(im);
	if (toupper(ox) > toupper(oy)) transpose_px(im);
}

/* a permutation of the axes (sample, column, row) named "x", "y", "z"
 * (iio.c:886-960) */
static void reorient3d(raw_image *im, const char *s)
{
	const size_t sb = codec_type_bytes(im->type);
	const int a = im->pd, b = im->w, c = im->h;   /* sizes along x, y, z */
	int v0, v1, v2;
	if (!strcmp(s, "xyz"))      { v0 = a; v1 = b; v2 = c; }
	else if (!strcmp(s, "yxz")) { v0 = b; v1 = a; v2 = c; }
	else if (!strcmp(s, "xzy")) { v0 = a; v1 = c; v2 = b; }
	else if (!strcmp(s, "zxy")) { v0 = c; v1 = a; v2 = b; }
	else if (!strcmp(s, "zyx")) { v0 = c; v1 = b; v2 = a; }
	else if (!strcmp(s, "yzx")) { v0 = b; v1 = c; v2 = a; }
	else { io_fail("unrecognized trasnposition string \"%s\"", s); return; }
	unsigned char *t = xalloc((size_t)a * b * c * sb);
	for (int k = 0; k < c; k++)
	for (int j = 0; j < b; j++)
	for (int i = 0; i < a; i++) {
		/* coordinate of each output axis: the input coordinate of the
		 * axis its letter names */
		const int co[3] = {i, j, k};
		const int o0 = co[s[0] - 'x'], o1 = co[s[1] - 'x'], o2 = co[s[2] - 'x'];
		memcpy(t + (((size_t)o2 * v1 + o1) * v0 + o0) * sb,
			im->data + (((size_t)k * b + j) * a + i) * sb, sb);
	}
	free(im->data);
	im->data = t;
	im->pd = v0;
	im->w = v1;
	im->h = v2;
}

/* Crop (or pad, for negative amounts) by the given numbers of columns and
 * rows (iio.c:832-870).  New samples are NaN for 4- and 8-byte types, zero
 * bytes otherwise, as iio fills them whatever the type.  iio's NaN is
 * 0.0f/0.0f evaluated at run time, the x86 default NaN: sign bit set. */
static void trim_px(raw_image *im, int left, int bottom, int right, int top)
{
	const size_t sb = codec_type_bytes(im->type), ps = px_bytes(im);
	const int w = im->w, h = im->h;
	const int nw = w - left - right, nh = h - bottom - top;
	if (nw <= 0 || nh <= 0)
		io_fail("image transformation leaves a %dx%d image", nw, nh);
	const size_t ns = (size_t)nw * nh * im->pd;
	unsigned char *t = xalloc(ns * sb);
	if (sb == 4) {
		const uint32_t nan4 = 0xffc00000u;
		for (size_t k = 0; k < ns; k++) memcpy(t + 4 * k, &nan4, 4);
	} else if (sb == 8) {
		const uint64_t nan8 = 0xfff8000000000000ull;
		for (size_t k = 0; k < ns; k++) memcpy(t + 8 * k, &nan8, 8);
	} else {
		memset(t, 0, ns * sb);
	}
	for (int j = 0; j < nh; j++)
	for (int i = 0; i < nw; i++) {
		const int x = i + left, y = j + top;
		if (x >= 0 && y >= 0 && x < w && y < h)
			memcpy(t + ((size_t)j * nw + i) * ps,
				im->data + ((size_t)y * w + x) * ps, ps);
	}
	free(im->data);
	im->data = t;
	im->w = nw;
	im->h = nh;
}

/* "flip=..." (iio.c:3810-3823): every test applies in turn */
static void trans_flip(raw_image *im, const char *s)
{
	if (!strcmp(s, "leftright")) flip_leftright(im);
	if (!strcmp(s, "topdown")) flip_topdown(im);
	if (!strcmp(s, "transpose")) transpose_px(im);
	if (!strcmp(s, "r270")) reorient(im, 'y', 'X');
	if (!strcmp(s, "r90")) reorient(im, 'Y', 'x');
	if (!strcmp(s, "r180")) reorient(im, 'X', 'Y');
	if (!strcmp(s, "posetrans")) reorient(im, 'Y', 'X');
	const size_t n = strlen(s);
	if (n == 2) reorient(im, (unsigned char)s[0], (unsigned char)s[1]);
	if (n == 3 && 3 * 'y' == tolower((unsigned char)s[0])
			+ tolower((unsigned char)s[1]) + tolower((unsigned char)s[2]))
		reorient3d(im, s);
}

static const char *npy_descr(int type)
{
	switch (type) {
	case T_F4: return "<f4";
	case T_F8: return "<f8";
	case T_U1: return "<u1";
	case T_U2: return "<u2";
	case T_U4: return "<u4";
	case T_U8: return "<u8";
	case T_I1: return "<i1";
	case T_I2: return "<i2";
	case T_I4: return "<i4";
	default: return "<i8";
	}
}

static void npy_header(FILE *f, const char *descr, int w, int h, int pd);

/* "pipe=COMMAND" (iio.c:3829-3852): the image goes to the command's stdin
 * as NPY in its sample type, and what the command writes on stdout is read
 * back, if the command succeeds */
static void trans_pipe(raw_image *im, const char *cmd)
{
	char in[] = "/tmp/iio_tmp_file_XXXXXX", out[] = "/tmp/iio_tmp_file_XXXXXX";
	int fi = mkstemp(in), fo = mkstemp(out);
	if (fi < 0 || fo < 0) io_fail("could not create tmp filename");
	close(fo);
	FILE *f = fdopen(fi, "wb");
	if (!f) io_fail("could not create tmp filename");
	npy_header(f, npy_descr(im->type), im->w, im->h, im->pd);
	fwrite(im->data, 1, (size_t)im->w * im->h * px_bytes(im), f);
	fclose(f);
	const size_t len = strlen(cmd) + strlen(in) + strlen(out) + 40;
	char *c = xalloc(len);
	snprintf(c, len, "(unset IIO_TRANS; %s) <%s >%s", cmd, in, out);
	fflush(NULL);
	if (!system(c)) {
		free(im->data);
		read_core(out, im);
	}
	free(c);
	remove(in);
	remove(out);
}

/* "OP=VALUE,OP=VALUE..." up to the first ']' (iio.c:3854-3884); tokens
 * without '=' are skipped */
static void trans_apply(raw_image *im, const char *f)
{
	size_t n = strcspn(f, "]");
	char *t = xalloc(n + 1), *o = xalloc(n + 1), *v = xalloc(n + 1);
	memcpy(t, f, n);
	t[n] = 0;
	for (char *tok = strtok(t, ","); tok; tok = strtok(NULL, ",")) {
		if (sscanf(tok, "%[^=]=%[^]]", o, v) != 2) continue;
		const int V = atoi(v), w = im->w, h = im->h;
		if (!strcmp(o, "flip")) trans_flip(im, v);
		else if (!strcmp(o, "pipe")) trans_pipe(im, v);
		else if (!strcmp(o, "x")) trim_px(im, V, 0, 0, 0);
		else if (!strcmp(o, "y")) trim_px(im, 0, 0, 0, V);
		else if (!strcmp(o, "w")) trim_px(im, 0, 0, w - V, 0);
		else if (!strcmp(o, "h")) trim_px(im, 0, h - V, 0, 0);
	}
	free(t);
	free(o);
	free(v);
}

/* "TRANS[list]:name" gives the colon, or NULL (iio.c:3800-3808) */
static const char *trans_prefix(const char *f)
{
	if (strncmp(f, "TRANS[", 6)) return NULL;
	const char *colon = strchr(f, ':');
	if (!colon || colon[-1] != ']') return NULL;
	return colon;
}

/* "zero:WxH[,pd]", "one:WxH[,pd]" (bytes) and "constant:V:WxH" (float),
 * iio.c:5603-5634; returns 0 for other names */
static int read_semantic(const char *name, raw_image *im)
{
	int s[2], pd = 1, one = 0;
	float value = 0;
	if (!strncmp(name, "zero:", 5) || (one = !strncmp(name, "one:", 4))) {
		const char *a = name + (one ? 4 : 5);
		if (sscanf(a, "%dx%d,%d", s, s + 1, &pd) != 3 &&
				sscanf(a, "%dx%d", s, s + 1) != 2)
			io_fail("bad semantical name \"%s\"", name);
		im->type = T_U1;
	} else if (!strncmp(name, "constant:", 9)) {
		if (sscanf(name + 9, "%g:%dx%d", &value, s, s + 1) != 3)
			io_fail("bad semantical name \"%s\"", name);
		im->type = T_F4;
	} else {
		return 0;
	}
	if (s[0] <= 0 || s[1] <= 0 || pd <= 0)
		io_fail("bad semantical name \"%s\"", name);
	const size_t ns = (size_t)s[0] * s[1] * pd;
	im->w = s[0]; im->h = s[1]; im->pd = pd;
	im->data = xalloc(ns * codec_type_bytes(im->type));
	if (im->type == T_U1)
		memset(im->data, one, ns);
	else
		for (size_t k = 0; k < ns; k++) memcpy(im->data + 4 * k, &value, 4);
	return 1;
}

/* iio's read_image (iio.c:5585-5706): semantic names first (no IIO_TRANS
 * for them), then "TRANS[...]:" names recursively, then files; IIO_TRANS
 * from the environment applies last, so a "TRANS[...]:" name gets it twice
 * (once inside), as in the reference */
static void read_image(const char *name, raw_image *im)
{
	if (read_semantic(name, im)) return;
	const char *colon = trans_prefix(name);
	if (colon) {
		read_image(colon + 1, im);
		trans_apply(im, name + 6);
	} else {
		read_core(name, im);
	}
	const char *t = getenv("IIO_TRANS");
	if (t && *t) trans_apply(im, t);
}

/* interleaved floats of an image as read (iio_convert_samples,
 * iio.c:980-1090: plain casts); frees the samples */
static float *to_float(raw_image *im)
{
	const size_t ns = (size_t)im->w * im->h * im->pd;
	const int sb = codec_type_bytes(im->type);
	float *px = xalloc(ns * sizeof(float));
	for (size_t k = 0; k < ns; k++)
		px[k] = sample_to_float(im->data + k * sb, im->type);
	free(im->data);
	im->data = NULL;
	return px;
}

/* planar float image, iio_read_image_float_split (iio.c:5763-5771) */
static float *read_split(const char *name, int *w, int *h, int *pd)
{
	raw_image im;
	read_image(name, &im);
	float *px = to_float(&im);
	size_t n = (size_t)im.w * im.h;
	float *out = xalloc(n * im.pd * sizeof(float));
	for (size_t i = 0; i < n; i++)
		for (int l = 0; l < im.pd; l++)
			out[n * l + i] = px[im.pd * i + l];
	free(px);
	*w = im.w; *h = im.h; *pd = im.pd;
	return out;
}

/* Gray of the first three samples of each pixel, computed in double and
 * stored in the sample type (iio_hacky_uncolorize[a], iio.c:1230-1300):
 * 8-bit, 16-bit (three samples only) and float images */
static void uncolorize(raw_image *im)
{
	const int t = im->type, pd = im->pd;
	if (!(t == T_U1 || t == T_F4 || (t == T_U2 && pd == 3)))
		io_fail(pd == 3 ? "uncolorize type not supported"
				: "uncolorizea type not supported");
	const size_t n = (size_t)im->w * im->h;
	const int sb = codec_type_bytes(t);
	unsigned char *g = xalloc(n * sb);
	for (size_t i = 0; i < n; i++) {
		const unsigned char *p = im->data + i * pd * sb;
		const double v = .299 * sample_to_float(p, t)
			+ .587 * sample_to_float(p + sb, t)
			+ .114 * sample_to_float(p + 2 * sb, t);
		if (t == T_U1) { g[i] = (uint8_t)v; }
		else if (t == T_U2) { uint16_t s = (uint16_t)v; memcpy(g + 2 * i, &s, 2); }
		else { float s = (float)v; memcpy(g + 4 * i, &s, 4); }
	}
	free(im->data);
	im->data = g;
	im->pd = 1;
}

/* scalar mask, iio_read_image_float (iio.c:6029-6048) */
static float *read_mask(const char *name, int *w, int *h)
{
	raw_image im;
	read_image(name, &im);
	if (im.pd == 3 || im.pd == 4) uncolorize(&im);
	if (im.pd != 1) io_fail("could not read image");
	*w = im.w; *h = im.h;
	return to_float(&im);
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
static void npy_header(FILE *f, const char *descr, int w, int h, int pd)
{
	char buf[1000];
	buf[0] = (char)0x93;
	memcpy(buf + 1, "NUMPY", 5);
	buf[6] = 1; buf[7] = 0;
	int n = 10;
	n += snprintf(buf + n, sizeof buf - n,
		"{'descr': '%s', 'fortran_order': False, 'shape': (%d, %d, %d)}",
		descr, h, w, pd);
	int m = ((n + 15 + 1) / 16) * 16;
	buf[8] = (char)((m - 10) % 256);
	buf[9] = (char)((m - 10) / 256);
	for (int i = n; i < m - 1; i++) buf[i] = ' ';
	buf[m - 1] = '\n';
	fwrite(buf, 1, m, f);
}

static void write_npy(const char *name, const float *px, int w, int h, int pd)
{
	FILE *f = open_out(name);
	npy_header(f, "<f4", w, h, pd);
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
		read_image(fa, &im);
		A = to_float(&im);
		if (im.w != w || im.h != h || im.pd != 3)
			return fprintf(stderr, "ERROR: bad A dimensions"
				"(%d %d %d) ((%d %d))\n", im.w, im.h, im.pd, w, h);
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
	read_image(fname, &im);
	float *px = to_float(&im);
	*w = im.w; *h = im.h; *pd = im.pd;
	return px;
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
