// Tile-shape autotuner of the integer search (rows per lane K, lanes per
// candidate row LC), keyed by (PU shape, search range).
#pragma once

struct TuneChoice {
    int K;    // candidate rows per lane (4 or 8; 16 for narrow shapes)
    int LC;   // lanes sharing one candidate row, 4 columns each
};

// Tuned choice when one exists, else the default for the shape.
TuneChoice xme_tune_lookup(int part, int search_range);
TuneChoice xme_tune_default(int part, int search_range);
void xme_tune_store(int part, int search_range, TuneChoice c);
// Forces a choice for every shape (used by the tuner while it measures).
void xme_tune_force(const TuneChoice *c);

int xme_part_index(int w, int h);
// At most 256 pixels (every shape up to 16 x 16, 8 x 32 and 32 x 8):
// eligible for 16 candidate rows per lane (the short-PU kernel variant).
bool xme_narrow_shape(int part);
int xme_pu_w(int part);
int xme_pu_h(int part);
const void *xme_fs_kernel_ptr(bool persistent, int kmax);
// Rows-per-lane bound of the search kernel used by new plans (4 or 8).
int xme_fs_kmax(void);
void xme_fs_set_kmax(int kmax);
// Search workgroups that share a CU's LDS (windows are split to fit), 1..4;
// XME_FS_WG_PER_CU overrides the default of 2.
int xme_fs_wg_per_cu(void);
