// irlprj.h -- internal types shared by the host layers of libirl_amd.
// The C declarations exported to callers live in include/irlprj.h; this file
// only adds the C++ pieces the host layers share.
#pragma once

#include <cstdio>
#include <string>
#include <vector>

#include "../../include/irl.h"
#include "../../include/irlprj.h"
#include "../../include/osemhooks.h"
#include "engine.h"

namespace irlamd {

// Message and log plumbing (the reference's vPrintMsg / vErrorHandler
// behaviour: level-filtered status messages, warnings to the log file).
void msg(int level, const char *fmt, ...);
void warn(const char *fmt, ...);
void set_msg_level(int level);
void set_msg_file(FILE *fp);
void set_log_file(FILE *fp);
void set_error_string(const std::string &s);
void set_last_error(int code);
int last_error();

// Reference-arithmetic host geometry.
int drf_row_index(const DrfTab_t *t, float fCFCR, float fPixelWidth, int iNumRotPixs, int iRow,
                  bool *skip);
PaddedDrf pad_drf_table(const DrfTab_t *t);
void view_trig(float angle, ViewTab &vt);

// Computed parallel-hole GRF tables (gendrf.c:221-427, IrlSetup.c:243-334).
void gen_grf_table(float fColThick, float fHoleDiam, float fGap, float fIntrinsic, float fMaxErr,
                   float fBinSize, float fDrfSpacing, float minDist, int NumDists, int NOffs,
                   Drf_t *tab);
void compute_grf_tab(DrfTab_t *t, float fPixelWidth, float fBinWidth, int NRot, float fMinDist,
                     float fMaxDist);

// DRFT / KRNL file readers (drftabio.c:150-275, krnlio.c:196-331).
DrfTab_t read_drf_tab(const char *fname, std::string *title, std::string *col);
void free_drf_tab_contents(DrfTab_t *t);

}  // namespace irlamd
