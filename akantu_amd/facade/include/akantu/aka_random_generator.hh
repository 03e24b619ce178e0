/**
 * aka_random_generator.hh -- the random parameters of the façade's input
 * files, drawn as the reference draws them: one std::default_random_engine
 * seeded with the run's seed (RandomGenerator, aka_random_generator.hh:72-96,
 * seeded in aka_common.cc:99-108), each value base + distribution(engine)
 * (RandomParameter::setValues, aka_random_generator.hh:230-234), with the
 * parameter order of the input file's `name = base dist [p0, p1]`
 * (parser_random.cc:71-83: uniform [min, max], weibull [scale, shape]).
 */
#pragma once

#include <cstdint>
#include <random>
#include <string>

#include "aka_common.hh"

namespace akantu {

/// n draws of `base + dist(p0, p1)` from `engine`, in order; dist is
/// "uniform" or "weibull"; throws on another name
void drawRandomValues(std::default_random_engine& engine, const std::string& dist, Real base,
                      Real p0, Real p1, Real* out, size_t n);

}  // namespace akantu

extern "C" {
/// The same from a fresh engine seeded with `seed` (the draws initFull makes
/// for the first random parameter of a run).  0 on success, -1 on an unknown
/// distribution.
int akantu_random_values(long seed, const char* dist, double base, double p0, double p1,
                         int64_t n, double* out);
}
