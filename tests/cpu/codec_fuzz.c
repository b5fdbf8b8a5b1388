/*
 * Robustness of the PNG / TIFF decoders (imscript_amd/csrc/sp_codecs.c)
 * on damaged files, used by the test-suite: every input file given is
 * decoded as is, then with random byte changes, truncations and chunk
 * swaps (seeded).  Built with AddressSanitizer by the tests, so any read or
 * write out of bounds aborts the run; a decoder may refuse a file, never
 * crash on it.
 *
 * usage: codec_fuzz SEED ROUNDS file...
 * prints "OK <decoded> <refused>"
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../imscript_amd/csrc/sp_codecs.h"

static unsigned rs;
static unsigned rnd(void) { rs = rs * 1103515245u + 12345u; return rs >> 8; }

static int decode(const unsigned char *b, size_t n, int *ok, int *bad)
{
	codec_image im;
	char err[256];
	int r = n >= 4 && b[0] == 0x89 ? codec_read_png(b, n, &im, err, sizeof err)
		: codec_read_tiff(b, n, &im, err, sizeof err);
	if (r) { (*bad)++; return 0; }
	(*ok)++;
	/* touch every sample */
	size_t ns = (size_t)im.w * im.h * im.pd * codec_type_bytes(im.type);
	unsigned s = 0;
	for (size_t k = 0; k < ns; k++) s += im.data[k];
	free(im.data);
	return (int)(s & 1);
}

int main(int argc, char **argv)
{
	if (argc < 4) { fprintf(stderr, "usage: codec_fuzz SEED ROUNDS file...\n"); return 2; }
	rs = (unsigned)atoi(argv[1]) * 2654435761u + 1;
	const int rounds = atoi(argv[2]);
	int ok = 0, bad = 0;
	for (int f = 3; f < argc; f++) {
		FILE *fp = fopen(argv[f], "rb");
		if (!fp) { perror(argv[f]); return 2; }
		fseek(fp, 0, SEEK_END);
		long n = ftell(fp);
		fseek(fp, 0, SEEK_SET);
		unsigned char *b = malloc(n), *m = malloc(n + 64);
		if (fread(b, 1, n, fp) != (size_t)n) return 2;
		fclose(fp);
		decode(b, n, &ok, &bad);
		for (int r = 0; r < rounds; r++) {
			size_t len = n;
			memcpy(m, b, n);
			const int kind = rnd() % 4;
			const int edits = 1 + rnd() % 8;
			for (int e = 0; e < edits; e++) {
				/* changes near the start hit headers and directories */
				size_t at = rnd() % 3 ? rnd() % (len < 512 ? len : 512) : rnd() % len;
				if (kind == 0) m[at] = (unsigned char)rnd();
				else if (kind == 1) m[at] ^= (unsigned char)(1u << (rnd() % 8));
				else if (kind == 2) m[at] = (unsigned char)(rnd() % 2 ? 0xff : 0);
			}
			if (kind == 3) len = 8 + rnd() % (len - 8);
			decode(m, len, &ok, &bad);
		}
		free(b);
		free(m);
	}
	printf("OK %d %d\n", ok, bad);
	return 0;
}
