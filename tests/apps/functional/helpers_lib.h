// A helper of the helpers program kept in a header of its own: the
// translator finds called functions in quoted includes too.
#pragma once

static inline double cube(double v) { return v * v * v; }
