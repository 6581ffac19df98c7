/*
 * mcx_media.h -- values of Config.mediabyte above 4: the per-voxel
 * property encodings of the reference (src/mcx_const.h:72-80).  Shared by
 * the C ABI header and the transport kernel.
 */
#ifndef MCX_AMD_MEDIA_H
#define MCX_AMD_MEDIA_H

/* bit 31 of every word is the detector mask */
#define MCX_MEDIA_ASGN_F2H     96   /* 2 words per voxel: half {mua, mus}, then (nvox words on) half {g, n} */
#define MCX_MEDIA_2LABEL_SPLIT 97   /* 2 words per voxel: split-voxel labels, interface point and normal */
#define MCX_MEDIA_2LABEL_MIX   98   /* byte label0, byte label1, short label1 fraction x 32767 */
#define MCX_MEDIA_LABEL_HALF   99   /* short: label (14 bits) + replaced property index (2 bits); half value */
#define MCX_MEDIA_AS_F2H       100  /* half {mua, mus} (converted from float input) */
#define MCX_MEDIA_MUA_FLOAT    101  /* float mua; mus, g, n of prop row 1 */
#define MCX_MEDIA_AS_HALF      102  /* half {mua, mus} */
#define MCX_MEDIA_ASGN_BYTE    103  /* bytes: mua, mus, g, n as fractions between prop rows 1 and 2 */
#define MCX_MEDIA_AS_SHORT     104  /* shorts: mua, mus as fractions between prop rows 1 and 2 */

#endif
