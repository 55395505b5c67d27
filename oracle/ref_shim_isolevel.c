/*
 * TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
 * reference's own src/isolevel.c (included from where it lies, never copied)
 * to expose the file-static nii_otsu (src/isolevel.c:141) to the parity
 * tests; setThreshold and the rest of the file are exported as they are.
 */
#include "isolevel.c"

int ref_nii_otsu(int *H, int nBin, int mode, int *dark, int *mid, int *bright)
{
	return nii_otsu(H, nBin, mode, dark, mid, bright);
}
