// Type aliases of the reference's interface (src/lible/types.hpp:9-33): fixed-size nested
// arrays (the second-derivative kernel returns arr2d<vec2d, 6, 6>) and the VectorMD aliases.
#pragma once

#include <array>
#include <cstddef>
#include <vector>

#include "vectormd.hpp"

namespace lible
{
    template <typename T, std::size_t d1, std::size_t d2>
    using arr2d = std::array<std::array<T, d2>, d1>;
    template <typename T, std::size_t d1, std::size_t d2, std::size_t d3>
    using arr3d = std::array<std::array<std::array<T, d3>, d2>, d1>;
    template <typename T, std::size_t d1, std::size_t d2, std::size_t d3, std::size_t d4>
    using arr4d = std::array<std::array<std::array<std::array<T, d4>, d3>, d2>, d1>;
    template <typename T>
    using vecvec = std::vector<std::vector<T>>;
}
