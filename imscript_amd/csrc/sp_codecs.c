/*
 * sp_codecs.c -- PNG and TIFF coding for the command lines (see
 * sp_codecs.h).  The reference reads and writes these formats through
 * libpng and libtiff (iio.c:1500-1556, 1668-1915, 3940-4081); this file
 * implements the parts of the file formats those paths use, on zlib:
 *
 *   PNG  (ISO/IEC 15948): all colour types and bit depths, Adam7
 *        interlacing, the five row filters, PLTE and tRNS, with the
 *        results of libpng's PACKING | EXPAND transforms: palettes become
 *        RGB(A), gray below 8 bits is scaled to 8 bits, tRNS becomes an
 *        alpha channel.  Written: 8- or 16-bit gray, gray+alpha, RGB, RGBA.
 *   TIFF (6.0): the first image, strips or tiles, chunky or planar,
 *        compression none, LZW, Deflate (8 and 32946) and PackBits,
 *        predictors 1, 2 (horizontal) and 3 (floating point, Adobe
 *        technical note 3), 1/2/4-bit unsigned samples unpacked to bytes,
 *        8-64-bit integers, 32/64-bit floats, complex formats split into
 *        two samples.  Written: float32, chunky, one strip, LZW or none.
 */
#include "sp_codecs.h"

#include <stdarg.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

static int cfail(char *err, size_t errn, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	if (err && errn) vsnprintf(err, errn, fmt, ap);
	va_end(ap);
	return -1;
}

int codec_type_bytes(int t)
{
	switch (t) {
	case T_U1: case T_I1: return 1;
	case T_U2: case T_I2: return 2;
	case T_F4: case T_U4: case T_I4: return 4;
	default: return 8;
	}
}

static uint32_t rd_be32(const unsigned char *p)
{
	return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
}

/* =============================================================== PNG */

static const unsigned char png_sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};

/* Adam7 passes 0-6: x0, y0, dx, dy; pass 7: the whole image (no
 * interlacing) */
static const int adam7[8][4] = {
	{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
	{0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}, {0, 0, 1, 1},
};

static int paeth(int a, int b, int c)
{
	int p = a + b - c, pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
	if (pa <= pb && pa <= pc) return a;
	return pb <= pc ? b : c;
}

/* undo the row filters of one pass in place: rows of 1 + rb bytes */
static int png_unfilter(unsigned char *d, int rows, size_t rb, int bpp)
{
	unsigned char *prev = NULL;
	for (int j = 0; j < rows; j++) {
		unsigned char *r = d + (size_t)j * (rb + 1), *x = r + 1;
		const int ft = r[0];
		for (size_t i = 0; i < rb; i++) {
			const int a = i >= (size_t)bpp ? x[i - bpp] : 0;
			const int b = prev ? prev[i] : 0;
			const int c = prev && i >= (size_t)bpp ? prev[i - bpp] : 0;
			switch (ft) {
			case 0: break;
			case 1: x[i] = (unsigned char)(x[i] + a); break;
			case 2: x[i] = (unsigned char)(x[i] + b); break;
			case 3: x[i] = (unsigned char)(x[i] + ((a + b) >> 1)); break;
			case 4: x[i] = (unsigned char)(x[i] + paeth(a, b, c)); break;
			default: return -1;
			}
		}
		prev = x;
	}
	return 0;
}

/* sample k (0-based, in channel order) of a packed row */
static unsigned png_sample(const unsigned char *x, size_t k, int depth)
{
	if (depth == 8) return x[k];
	if (depth == 16) return (unsigned)x[2 * k] << 8 | x[2 * k + 1];
	const size_t bit = k * depth;
	return (x[bit >> 3] >> (8 - depth - (bit & 7))) & ((1u << depth) - 1);
}

int codec_read_png(const unsigned char *buf, size_t n, codec_image *im,
		char *err, size_t errn)
{
	memset(im, 0, sizeof *im);
	if (n < 8 || memcmp(buf, png_sig, 8)) return cfail(err, errn, "not a PNG file");
	uint32_t w = 0, h = 0;
	int depth = 0, ctype = -1, inter = 0, npal = 0, ntrns = 0, have_trns = 0;
	unsigned char pal[256][3], trns[256];
	unsigned trns_gray = 0, trns_rgb[3] = {0, 0, 0};
	unsigned char *idat = NULL;
	size_t nidat = 0, pos = 8;
	int seen_end = 0;
	while (pos + 12 <= n) {
		const uint32_t len = rd_be32(buf + pos);
		const unsigned char *type = buf + pos + 4, *data = buf + pos + 8;
		if (len > n || pos + 12 + len > n) break;
		if (crc32(crc32(0, NULL, 0), type, len + 4) != rd_be32(data + len)) {
			free(idat);
			return cfail(err, errn, "png error (chunk CRC)");
		}
		if (!memcmp(type, "IHDR", 4) && len >= 13) {
			w = rd_be32(data);
			h = rd_be32(data + 4);
			depth = data[8];
			ctype = data[9];
			inter = data[12];
		} else if (!memcmp(type, "PLTE", 4)) {
			npal = (int)(len / 3 > 256 ? 256 : len / 3);
			for (int k = 0; k < npal; k++) memcpy(pal[k], data + 3 * k, 3);
		} else if (!memcmp(type, "tRNS", 4)) {
			have_trns = 1;
			if (ctype == 3) {
				ntrns = (int)(len > 256 ? 256 : len);
				memcpy(trns, data, (size_t)ntrns);
			} else if (ctype == 0 && len >= 2) {
				trns_gray = (unsigned)data[0] << 8 | data[1];
			} else if (ctype == 2 && len >= 6) {
				for (int k = 0; k < 3; k++)
					trns_rgb[k] = (unsigned)data[2 * k] << 8 | data[2 * k + 1];
			} else {
				have_trns = 0;
			}
		} else if (!memcmp(type, "IDAT", 4)) {
			unsigned char *t = realloc(idat, nidat + len + 1);
			if (!t) { free(idat); return cfail(err, errn, "out of memory"); }
			idat = t;
			memcpy(idat + nidat, data, len);
			nidat += len;
		} else if (!memcmp(type, "IEND", 4)) {
			seen_end = 1;
			break;
		}
		pos += 12 + len;
	}
	(void)seen_end;
	int ch;
	switch (ctype) {
	case 0: ch = 1; break;
	case 2: ch = 3; break;
	case 3: ch = 1; break;
	case 4: ch = 2; break;
	case 6: ch = 4; break;
	default: free(idat); return cfail(err, errn, "png error (color type %d)", ctype);
	}
	const int ok_depth = ctype == 0 ? (depth == 1 || depth == 2 || depth == 4 ||
			depth == 8 || depth == 16) : ctype == 3 ? (depth == 1 || depth == 2 ||
			depth == 4 || depth == 8) : (depth == 8 || depth == 16);
	if (!w || !h || !ok_depth || inter > 1 || !idat || (ctype == 3 && !npal)) {
		free(idat);
		return cfail(err, errn, "png error (header)");
	}
	if ((uint64_t)w * h > (1u << 30) || (uint64_t)w * h * ch * (depth > 8 ? 2 : 1) > (1ull << 31)) {
		free(idat);
		return cfail(err, errn, "png too large");
	}
	const int bits = depth * ch, bpp = bits < 8 ? 1 : bits / 8;
	/* passes 0-6 of Adam7, or the whole image as pass 7 */
	const int p0 = inter ? 0 : 7, p1 = inter ? 7 : 8;
	size_t total = 0;
	for (int p = p0; p < p1; p++) {
		const int *a = adam7[p];
		const size_t pw = w > (uint32_t)a[0] ? (w - a[0] + a[2] - 1) / a[2] : 0;
		const size_t ph = h > (uint32_t)a[1] ? (h - a[1] + a[3] - 1) / a[3] : 0;
		if (pw && ph) total += ph * (1 + (pw * bits + 7) / 8);
	}
	unsigned char *raw = malloc(total ? total : 1);
	uLongf got = (uLongf)total;
	const int zr = raw ? uncompress(raw, &got, idat, (uLong)nidat) : Z_MEM_ERROR;
	free(idat);
	if (zr != Z_OK && zr != Z_BUF_ERROR) { free(raw); return cfail(err, errn, "png error (zlib %d)", zr); }
	if (got < total) { free(raw); return cfail(err, errn, "png error (short image data)"); }
	/* output channels and depth after PACKING | EXPAND */
	const int och = ctype == 3 ? (have_trns ? 4 : 3)
		: (ctype == 0 || ctype == 2) && have_trns ? ch + 1 : ch;
	const int o16 = depth == 16;
	const int ob = o16 ? 2 : 1;
	unsigned char *out = malloc((size_t)w * h * och * ob);
	if (!out) { free(raw); return cfail(err, errn, "out of memory"); }
	const unsigned omax = o16 ? 65535u : 255u;
	size_t off = 0;
	for (int p = p0; p < p1; p++) {
		const int *a = adam7[p];
		const size_t pw = w > (uint32_t)a[0] ? (w - a[0] + a[2] - 1) / a[2] : 0;
		const size_t ph = h > (uint32_t)a[1] ? (h - a[1] + a[3] - 1) / a[3] : 0;
		if (!pw || !ph) continue;
		const size_t rb = (pw * bits + 7) / 8;
		if (png_unfilter(raw + off, (int)ph, rb, bpp)) {
			free(raw); free(out);
			return cfail(err, errn, "png error (filter type)");
		}
		for (size_t j = 0; j < ph; j++) {
			const unsigned char *x = raw + off + j * (rb + 1) + 1;
			const size_t y = a[1] + j * a[3];
			for (size_t i = 0; i < pw; i++) {
				const size_t xi = a[0] + i * a[2];
				unsigned v[4] = {0, 0, 0, 0};
				int nv = 0;
				if (ctype == 3) {
					const unsigned idx = png_sample(x, i, depth);
					for (int k = 0; k < 3; k++) v[k] = idx < (unsigned)npal ? pal[idx][k] : 0;
					nv = 3;
					if (have_trns) v[nv++] = (int)idx < ntrns ? trns[idx] : 255;
				} else {
					for (int k = 0; k < ch; k++) v[k] = png_sample(x, i * ch + k, depth);
					nv = ch;
					if (have_trns && ctype == 0)
						v[nv++] = v[0] == trns_gray ? 0 : omax;
					if (have_trns && ctype == 2)
						v[nv++] = v[0] == trns_rgb[0] && v[1] == trns_rgb[1] &&
							v[2] == trns_rgb[2] ? 0 : omax;
					if (ctype == 0 && depth < 8)
						v[0] *= 255u / ((1u << depth) - 1);
				}
				unsigned char *o = out + ((y * w + xi) * och) * ob;
				for (int k = 0; k < nv; k++) {
					if (o16) {
						const uint16_t s = (uint16_t)v[k];
						memcpy(o + 2 * k, &s, 2);
					} else {
						o[k] = (unsigned char)v[k];
					}
				}
			}
		}
		off += ph * (rb + 1);
	}
	free(raw);
	im->w = (int)w;
	im->h = (int)h;
	im->pd = och;
	im->type = o16 ? T_U2 : T_U1;
	im->data = out;
	return 0;
}

static void png_chunk(FILE *f, const char *type, const unsigned char *d, uint32_t len)
{
	unsigned char hdr[8] = {(unsigned char)(len >> 24), (unsigned char)(len >> 16),
		(unsigned char)(len >> 8), (unsigned char)len, 0, 0, 0, 0};
	memcpy(hdr + 4, type, 4);
	uLong c = crc32(crc32(0, NULL, 0), (const Bytef *)type, 4);
	if (len) c = crc32(c, d, len);
	unsigned char tail[4] = {(unsigned char)(c >> 24), (unsigned char)(c >> 16),
		(unsigned char)(c >> 8), (unsigned char)c};
	fwrite(hdr, 1, 8, f);
	if (len) fwrite(d, 1, len, f);
	fwrite(tail, 1, 4, f);
}

int codec_write_png(FILE *f, const unsigned char *px8, const unsigned short *px16,
		int w, int h, int pd)
{
	static const int ctypes[5] = {-1, 0, 4, 2, 6};
	if (pd < 1 || pd > 4 || w < 1 || h < 1) return -1;
	const int depth = px16 ? 16 : 8, ob = depth / 8;
	const size_t rb = (size_t)w * pd * ob, nraw = (size_t)h * (rb + 1);
	unsigned char *raw = malloc(nraw);
	uLongf nz = compressBound((uLong)nraw);
	unsigned char *z = malloc(nz);
	if (!raw || !z) { free(raw); free(z); return -1; }
	for (int j = 0; j < h; j++) {
		unsigned char *r = raw + (size_t)j * (rb + 1);
		r[0] = 0;
		for (size_t k = 0; k < (size_t)w * pd; k++) {
			const size_t q = (size_t)j * w * pd + k;
			if (px16) {
				r[1 + 2 * k] = (unsigned char)(px16[q] >> 8);
				r[2 + 2 * k] = (unsigned char)px16[q];
			} else {
				r[1 + k] = px8[q];
			}
		}
	}
	const int zr = compress2(z, &nz, raw, (uLong)nraw, 6);
	free(raw);
	if (zr != Z_OK) { free(z); return -1; }
	unsigned char ihdr[13] = {(unsigned char)(w >> 24), (unsigned char)(w >> 16),
		(unsigned char)(w >> 8), (unsigned char)w, (unsigned char)(h >> 24),
		(unsigned char)(h >> 16), (unsigned char)(h >> 8), (unsigned char)h,
		(unsigned char)depth, (unsigned char)ctypes[pd], 0, 0, 0};
	fwrite(png_sig, 1, 8, f);
	png_chunk(f, "IHDR", ihdr, 13);
	for (size_t o = 0; o < nz; o += 1u << 20) {
		const size_t len = nz - o < (1u << 20) ? nz - o : (1u << 20);
		png_chunk(f, "IDAT", z + o, (uint32_t)len);
	}
	png_chunk(f, "IEND", NULL, 0);
	free(z);
	return ferror(f) ? -1 : 0;
}

/* =============================================================== LZW */

long codec_lzw_decode(const unsigned char *in, size_t n, unsigned char *out,
		size_t cap)
{
	static __thread uint16_t prefix[4096], len[4096];
	static __thread unsigned char suffix[4096], first[4096];
	for (int i = 0; i < 256; i++) {
		prefix[i] = 0xffff;
		suffix[i] = first[i] = (unsigned char)i;
		len[i] = 1;
	}
	size_t bit = 0, o = 0;
	const size_t nbits = n * 8;
	int width = 9, next = 258, old = -1;
	while (bit + width <= nbits && o < cap) {
		unsigned code = 0;
		for (int k = 0; k < width; k++, bit++)
			code = code << 1 | ((in[bit >> 3] >> (7 - (bit & 7))) & 1);
		if (code == 257) break;
		if (code == 256) {
			next = 258;
			width = 9;
			old = -1;
			continue;
		}
		if (old < 0) {
			if (code > 255) return -1;
			out[o++] = (unsigned char)code;
			old = (int)code;
			continue;
		}
		unsigned src;
		int extra;
		if ((int)code < next) { src = code; extra = 0; }
		else if ((int)code == next) { src = (unsigned)old; extra = 1; }
		else return -1;
		const size_t L = len[src];
		size_t room = cap - o, keep = L < room ? L : room;
		unsigned p = src;
		for (size_t k = L; k-- > 0;) {
			if (k < keep) out[o + k] = suffix[p];
			p = prefix[p];
		}
		o += keep;
		if (extra && o < cap) out[o++] = first[old];
		if (next < 4096) {
			prefix[next] = (uint16_t)old;
			suffix[next] = extra ? first[old] : first[code];
			first[next] = first[old];
			len[next] = (uint16_t)(len[old] + 1);
			next++;
		}
		if (next >= (1 << width) - 1 && width < 12) width++;
		old = (int)code;
	}
	return (long)o;
}

typedef struct { unsigned char *b; size_t n; uint32_t acc; int nacc; } bitw;

static void bw_put(bitw *w, unsigned code, int width)
{
	w->acc = w->acc << width | code;
	w->nacc += width;
	while (w->nacc >= 8) {
		w->b[w->n++] = (unsigned char)(w->acc >> (w->nacc - 8));
		w->nacc -= 8;
	}
	w->acc &= (1u << w->nacc) - 1;
}

long codec_lzw_encode(const unsigned char *in, size_t n, unsigned char **out)
{
	enum { HSIZE = 16411 };   /* prime above 4 x 4096 */
	uint32_t *hkey = malloc(HSIZE * sizeof *hkey);
	uint16_t *hcode = malloc(HSIZE * sizeof *hcode);
	bitw w = {malloc(2 * n + 64), 0, 0, 0};
	if (!hkey || !hcode || !w.b) { free(hkey); free(hcode); free(w.b); return -1; }
	memset(hkey, 0xff, HSIZE * sizeof *hkey);
	int width = 9, next = 258;
	bw_put(&w, 256, width);
	if (n) {
		unsigned ent = in[0];
		for (size_t i = 1; i <= n; i++) {
			if (i < n) {
				const unsigned c = in[i];
				const uint32_t key = ent << 8 | c;
				size_t hp = key % HSIZE;
				while (hkey[hp] != 0xffffffffu && hkey[hp] != key) hp = (hp + 1) % HSIZE;
				if (hkey[hp] == key) { ent = hcode[hp]; continue; }
				bw_put(&w, ent, width);
				hkey[hp] = key;
				hcode[hp] = (uint16_t)next;
				ent = c;
			} else {
				bw_put(&w, ent, width);
			}
			next++;
			if (next == 4094) {
				bw_put(&w, 256, width);
				memset(hkey, 0xff, HSIZE * sizeof *hkey);
				next = 258;
				width = 9;
			} else if (next > (1 << width) - 1) {
				width++;
			}
		}
	}
	bw_put(&w, 257, width);
	if (w.nacc) bw_put(&w, 0, 8 - w.nacc);
	free(hkey);
	free(hcode);
	*out = w.b;
	return (long)w.n;
}

/* ============================================================== TIFF */

typedef struct {
	const unsigned char *b;
	size_t n;
	int be;
} tbuf;

static uint32_t t16(const tbuf *t, size_t o)
{
	if (o + 2 > t->n) return 0;
	const unsigned char *p = t->b + o;
	return t->be ? (uint32_t)p[0] << 8 | p[1] : (uint32_t)p[1] << 8 | p[0];
}

static uint32_t t32(const tbuf *t, size_t o)
{
	if (o + 4 > t->n) return 0;
	const unsigned char *p = t->b + o;
	return t->be ? rd_be32(p)
		: (uint32_t)p[3] << 24 | (uint32_t)p[2] << 16 | (uint32_t)p[1] << 8 | p[0];
}

/* the values of one IFD entry as 32-bit integers (BYTE, SHORT, LONG) */
static uint32_t *tag_values(const tbuf *t, size_t e, uint32_t *count)
{
	const uint32_t type = t16(t, e + 2), cnt = t32(t, e + 4);
	const int sz = type == 3 || type == 8 ? 2 : type == 4 || type == 9 || type == 13 ? 4
		: type == 1 || type == 6 || type == 7 ? 1 : 0;
	if (!sz || cnt == 0 || cnt > (1u << 28)) return NULL;
	const size_t bytes = (size_t)sz * cnt;
	const size_t base = bytes <= 4 ? e + 8 : t32(t, e + 8);
	if (base + bytes > t->n) return NULL;
	uint32_t *v = malloc(cnt * sizeof *v);
	if (!v) return NULL;
	for (uint32_t k = 0; k < cnt; k++)
		v[k] = sz == 1 ? t->b[base + k] : sz == 2 ? t16(t, base + 2 * k) : t32(t, base + 4 * k);
	*count = cnt;
	return v;
}

static long packbits_decode(const unsigned char *in, size_t n, unsigned char *out,
		size_t cap)
{
	size_t i = 0, o = 0;
	while (i < n && o < cap) {
		const int c = (signed char)in[i++];
		if (c >= 0) {
			for (int k = 0; k <= c && i < n && o < cap; k++) out[o++] = in[i++];
		} else if (c != -128) {
			if (i >= n) break;
			const unsigned char v = in[i++];
			for (int k = 0; k < 1 - c && o < cap; k++) out[o++] = v;
		}
	}
	return (long)o;
}

static long inflate_into(const unsigned char *in, size_t n, unsigned char *out, size_t cap)
{
	z_stream z;
	memset(&z, 0, sizeof z);
	if (inflateInit(&z) != Z_OK) return -1;
	z.next_in = (Bytef *)in;
	z.avail_in = (uInt)n;
	z.next_out = out;
	z.avail_out = (uInt)cap;
	const int r = inflate(&z, Z_FINISH);
	const long got = (long)(cap - z.avail_out);
	inflateEnd(&z);
	return r == Z_STREAM_END || r == Z_OK || r == Z_BUF_ERROR ? got : -1;
}

/* undo predictor 2 on one row of ns native samples of nb bytes, stride sp */
static void hor_acc(unsigned char *r, size_t ns, int nb, int sp)
{
	for (size_t i = (size_t)sp; i < ns; i++) {
		switch (nb) {
		case 1: r[i] = (unsigned char)(r[i] + r[i - sp]); break;
		case 2: { uint16_t a, b; memcpy(&a, r + 2 * i, 2); memcpy(&b, r + 2 * (i - sp), 2);
			a = (uint16_t)(a + b); memcpy(r + 2 * i, &a, 2); break; }
		case 4: { uint32_t a, b; memcpy(&a, r + 4 * i, 4); memcpy(&b, r + 4 * (i - sp), 4);
			a += b; memcpy(r + 4 * i, &a, 4); break; }
		default: { uint64_t a, b; memcpy(&a, r + 8 * i, 8); memcpy(&b, r + 8 * (i - sp), 8);
			a += b; memcpy(r + 8 * i, &a, 8); break; }
		}
	}
}

/* undo predictor 3 on one row of ns samples of nb bytes, stride sp: byte
 * differences, then byte planes (most significant first) back to native
 * little-endian samples */
static int fp_acc(unsigned char *r, size_t ns, int nb, int sp)
{
	const size_t cc = ns * nb;
	for (size_t i = (size_t)sp; i < cc; i++) r[i] = (unsigned char)(r[i] + r[i - sp]);
	unsigned char *tmp = malloc(cc ? cc : 1);
	if (!tmp) return -1;
	memcpy(tmp, r, cc);
	for (size_t s = 0; s < ns; s++)
		for (int k = 0; k < nb; k++) r[nb * s + k] = tmp[(size_t)(nb - 1 - k) * ns + s];
	free(tmp);
	return 0;
}

int codec_read_tiff(const unsigned char *buf, size_t n, codec_image *im,
		char *err, size_t errn)
{
	memset(im, 0, sizeof *im);
	if (n < 8) return cfail(err, errn, "not a TIFF file");
	tbuf t = {buf, n, buf[0] == 'M'};
	if (!((buf[0] == 'I' && buf[1] == 'I') || (buf[0] == 'M' && buf[1] == 'M')))
		return cfail(err, errn, "not a TIFF file");
	if (t16(&t, 2) == 43) return cfail(err, errn, "BigTIFF is not supported");
	if (t16(&t, 2) != 42) return cfail(err, errn, "not a TIFF file");
	const size_t ifd = t32(&t, 4);
	const uint32_t nent = t16(&t, ifd);
	if (!nent || ifd + 2 + 12 * (size_t)nent > n) return cfail(err, errn, "bad TIFF directory");
	uint32_t w = 0, h = 0, spp = 1, bps = 1, fmt = 1, comp = 1, planar = 1, pred = 1;
	uint32_t rps = 0xffffffffu, tw = 0, tl = 0;
	uint32_t *offs = NULL, *cnts = NULL, noffs = 0, ncnts = 0;
	int have_w = 0, have_h = 0;
	for (uint32_t k = 0; k < nent; k++) {
		const size_t e = ifd + 2 + 12 * (size_t)k;
		const uint32_t tag = t16(&t, e);
		uint32_t cnt = 0, *v = NULL;
		switch (tag) {
		case 256: case 257: case 258: case 259: case 277: case 278: case 284:
		case 317: case 322: case 323: case 339:
			v = tag_values(&t, e, &cnt);
			if (!v) break;
			if (tag == 256) { w = v[0]; have_w = 1; }
			if (tag == 257) { h = v[0]; have_h = 1; }
			if (tag == 258) bps = v[0];
			if (tag == 259) comp = v[0];
			if (tag == 277) spp = v[0];
			if (tag == 278) rps = v[0];
			if (tag == 284) planar = v[0];
			if (tag == 317) pred = v[0];
			if (tag == 322) tw = v[0];
			if (tag == 323) tl = v[0];
			if (tag == 339) fmt = v[0];
			free(v);
			break;
		case 273: case 324:
			free(offs);
			offs = tag_values(&t, e, &noffs);
			break;
		case 279: case 325:
			free(cnts);
			cnts = tag_values(&t, e, &ncnts);
			break;
		}
	}
	int rc = 0;
	unsigned char *data = NULL, *chunk = NULL;
#define TFAIL(...) do { rc = cfail(err, errn, __VA_ARGS__); goto done; } while (0)
	if (!have_w || !have_h) TFAIL("can not read tiff of unknown size");
	if (!w || !h || (uint64_t)w * h > (1u << 30)) TFAIL("bad TIFF size");
	if (!offs || !cnts || noffs != ncnts) TFAIL("TIFF without image data");
	if (fmt == 5 || fmt == 6) { spp *= 2; bps /= 2; fmt = fmt == 5 ? 2 : 3; }
	int type;
	if (fmt == 1) {
		if (bps == 1 || bps == 2 || bps == 4 || bps == 8) type = T_U1;
		else if (bps == 16) type = T_U2;
		else if (bps == 32) type = T_U4;
		else TFAIL("unrecognized UINT type of size %u bits", bps);
	} else if (fmt == 2) {
		if (bps == 8) type = T_I1;
		else if (bps == 16) type = T_I2;
		else if (bps == 32) type = T_I4;
		else TFAIL("unrecognized INT type of size %u bits", bps);
	} else if (fmt == 3) {
		if (bps == 32) type = T_F4;
		else if (bps == 64) type = T_F8;
		else TFAIL("unrecognized FLOAT type of size %u bits", bps);
	} else {
		TFAIL("unrecognized tiff sample format %u (see tiff.h)", fmt);
	}
	if (spp < 1 || spp > 64) TFAIL("bad TIFF samples per pixel %u", spp);
	/* sizes a damaged file could claim are refused before allocating */
	if ((uint64_t)w * h * spp > (1ull << 31)) TFAIL("TIFF too large");
	if (tw || tl) {
		if (!tw || !tl || tw > 65536 || tl > 65536 || (uint64_t)tw * tl > (1ull << 26))
			TFAIL("bad TIFF tile size %ux%u", tw, tl);
	}
	if (comp != 1 && comp != 5 && comp != 8 && comp != 32946 && comp != 32773)
		TFAIL("unsupported TIFF compression %u", comp);
	if (pred != 1 && pred != 2 && pred != 3) TFAIL("unsupported TIFF predictor %u", pred);
	if (pred == 3 && fmt != 3) TFAIL("floating-point predictor on integer samples");
	if (bps < 8 && pred != 1) TFAIL("predictor on packed samples");
	const int nb = codec_type_bytes(type);
	const int tiled = tw && tl;
	const int sep = planar == 2 && spp > 1;
	const uint32_t cspp = sep ? 1 : spp;              /* samples per pixel per chunk */
	const uint32_t cw = tiled ? tw : w;                /* chunk width */
	uint32_t ch = tiled ? tl : (rps == 0 || rps > h ? h : rps);
	const size_t rowb = ((size_t)cw * cspp * bps + 7) / 8;
	const size_t need = rowb * ch;
	if (need > (1ull << 31)) TFAIL("TIFF strip or tile too large");
	const uint32_t across = tiled ? (w + tw - 1) / tw : 1;
	const uint32_t down = (h + ch - 1) / ch;
	const uint32_t per_plane = across * down;
	if (noffs < per_plane * (sep ? spp : 1)) TFAIL("TIFF with too few strips or tiles");
	data = calloc((size_t)w * h * spp, (size_t)nb);
	chunk = malloc(need ? need : 1);
	if (!data || !chunk) TFAIL("out of memory");
	for (uint32_t c = 0; c < per_plane * (sep ? spp : 1); c++) {
		const uint32_t plane = c / per_plane, q = c % per_plane;
		const uint32_t x0 = tiled ? (q % across) * tw : 0, y0 = (tiled ? q / across : q) * ch;
		const size_t so = offs[c], sn = cnts[c];
		if (so > n || sn > n - so) TFAIL("TIFF strip or tile outside the file");
		const unsigned char *src = buf + so;
		memset(chunk, 0, need);
		long got;
		if (comp == 1) { got = (long)(sn < need ? sn : need); memcpy(chunk, src, (size_t)got); }
		else if (comp == 5) {
			if (sn >= 2 && src[0] == 0 && (src[1] & 1)) TFAIL("old-style LZW is not supported");
			got = codec_lzw_decode(src, sn, chunk, need);
		}
		else if (comp == 32773) got = packbits_decode(src, sn, chunk, need);
		else got = inflate_into(src, sn, chunk, need);
		if (got < 0) TFAIL("error decoding TIFF data");
		const uint32_t rows = tiled ? ch : (h - y0 < ch ? h - y0 : ch);
		if ((size_t)got < rowb * rows) TFAIL("error read tiff row %u/%u", y0, h);
		for (uint32_t r = 0; r < rows; r++) {
			unsigned char *row = chunk + r * rowb;
			const size_t ns = (size_t)cw * cspp;
			if (bps >= 8) {
				if (pred == 3) {
					if (fp_acc(row, ns, nb, (int)cspp)) TFAIL("out of memory");
				} else {
					if (t.be && nb > 1)
						for (size_t s = 0; s < ns; s++)
							for (int k = 0; k < nb / 2; k++) {
								unsigned char x = row[s * nb + k];
								row[s * nb + k] = row[s * nb + nb - 1 - k];
								row[s * nb + nb - 1 - k] = x;
							}
					if (pred == 2) hor_acc(row, ns, nb, (int)cspp);
				}
			}
			const uint32_t y = y0 + r;
			if (y >= h) break;
			for (uint32_t i = 0; i < cw && x0 + i < w; i++)
				for (uint32_t l = 0; l < cspp; l++) {
					const size_t si = (size_t)i * cspp + l;
					const size_t dst = (((size_t)y * w + x0 + i) * spp + (sep ? plane : l)) * nb;
					if (bps >= 8) {
						memcpy(data + dst, row + si * nb, (size_t)nb);
					} else {
						const size_t bit = si * bps;
						data[dst] = (row[bit >> 3] >> (8 - bps - (bit & 7))) & ((1u << bps) - 1);
					}
				}
		}
	}
	im->w = (int)w;
	im->h = (int)h;
	im->pd = (int)spp;
	im->type = type;
	im->data = data;
	data = NULL;
done:
#undef TFAIL
	free(data);
	free(chunk);
	free(offs);
	free(cnts);
	return rc;
}

/* little-endian IFD entry */
static void ifd_entry(unsigned char *p, int tag, int type, uint32_t count, uint32_t value)
{
	p[0] = (unsigned char)tag; p[1] = (unsigned char)(tag >> 8);
	p[2] = (unsigned char)type; p[3] = 0;
	for (int k = 0; k < 4; k++) p[4 + k] = (unsigned char)(count >> (8 * k));
	if (type == 3 && count == 1) {
		p[8] = (unsigned char)value; p[9] = (unsigned char)(value >> 8); p[10] = p[11] = 0;
	} else {
		for (int k = 0; k < 4; k++) p[8 + k] = (unsigned char)(value >> (8 * k));
	}
}

int codec_write_tiff(FILE *f, const float *px, int w, int h, int pd, int lzw)
{
	if (w < 1 || h < 1 || pd < 1 || pd > 64) return -1;
	const size_t nraw = (size_t)w * h * pd * 4;
	unsigned char *strip = NULL;
	long nstrip;
	if (lzw) {
		nstrip = codec_lzw_encode((const unsigned char *)px, nraw, &strip);
		if (nstrip < 0) return -1;
	} else {
		nstrip = (long)nraw;
	}
	/* header | strip | pd-sized arrays (bps, sampleformat) | IFD */
	const uint32_t strip_off = 8;
	uint32_t arr_off = strip_off + (uint32_t)nstrip;
	arr_off += arr_off & 1;
	const uint32_t bps_off = arr_off, fmt_off = arr_off + 2 * pd;
	uint32_t ifd_off = fmt_off + 2 * pd;
	ifd_off += ifd_off & 1;
	unsigned char hdr[8] = {'I', 'I', 42, 0, (unsigned char)ifd_off,
		(unsigned char)(ifd_off >> 8), (unsigned char)(ifd_off >> 16),
		(unsigned char)(ifd_off >> 24)};
	fwrite(hdr, 1, 8, f);
	fwrite(lzw ? strip : (const unsigned char *)px, 1, (size_t)nstrip, f);
	long pos = 8 + nstrip;
	if (pos & 1) { fputc(0, f); pos++; }
	for (int k = 0; k < pd; k++) { fputc(32, f); fputc(0, f); }
	for (int k = 0; k < pd; k++) { fputc(3, f); fputc(0, f); }
	pos += 4 * pd;
	if (pos & 1) { fputc(0, f); pos++; }
	unsigned char e[13][12];
	int ne = 0;
	ifd_entry(e[ne++], 256, 4, 1, (uint32_t)w);
	ifd_entry(e[ne++], 257, 4, 1, (uint32_t)h);
	/* SHORT arrays of up to two values sit in the entry itself */
	ifd_entry(e[ne++], 258, 3, (uint32_t)pd, pd == 1 ? 32 : pd == 2 ? 32 | 32u << 16 : bps_off);
	ifd_entry(e[ne++], 259, 3, 1, lzw ? 5 : 1);
	ifd_entry(e[ne++], 262, 3, 1, pd == 3 || pd == 4 ? 2 : 1);
	ifd_entry(e[ne++], 273, 4, 1, strip_off);
	ifd_entry(e[ne++], 277, 3, 1, (uint32_t)pd);
	ifd_entry(e[ne++], 278, 4, 1, (uint32_t)h);
	ifd_entry(e[ne++], 279, 4, 1, (uint32_t)nstrip);
	ifd_entry(e[ne++], 284, 3, 1, 1);
	if (pd == 4) ifd_entry(e[ne++], 338, 3, 1, 2);
	ifd_entry(e[ne++], 339, 3, (uint32_t)pd, pd == 1 ? 3 : pd == 2 ? 3 | 3u << 16 : fmt_off);
	unsigned char cnt[2] = {(unsigned char)ne, 0}, next[4] = {0, 0, 0, 0};
	fwrite(cnt, 1, 2, f);
	fwrite(e, 12, (size_t)ne, f);
	fwrite(next, 1, 4, f);
	free(strip);
	return ferror(f) ? -1 : 0;
}
