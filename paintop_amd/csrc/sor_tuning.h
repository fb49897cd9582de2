// sor_tuning.h -- half-sweeps per launch (K) and sweep-tile rows (hb) by the
// size of a call, written by the tuner (python -m paintop_amd.tune) from timed
// plain and learned solves of the seamless-clone generator
// (device: AMD Instinct (gfx950)). Work W of a call = the sum over its images of
// channels x the area of the interior's bounding box; the first row with
// W <= w_max applies. Plain and learned calls share the table.
#pragma once

struct PhSorTune {
  double w_max;
  int k, hb;
};

static const PhSorTune PH_SOR_TUNING[] = {
    {1.749e+07, 8, 24},
    {1.000e+300, 16, 40},
};
