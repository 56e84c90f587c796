// Row-major N-dimensional array, the return type of the lible::ints integral functions.
// Same contract as the reference's VectorMD (src/lible/vectormd.hpp:24-27): contiguous
// row-major storage, operator()(i, j, ...) element access, dim<k>(), size(), iteration over all
// elements, memptr().
#pragma once

#include <algorithm>
#include <array>
#include <cstddef>
#include <vector>

namespace lible
{
    template <typename T>
    struct Fill
    {
        explicit Fill(T v) : val_(v) {}
        const T val_;
    };

    template <typename T, std::size_t N>
    class VectorMD
    {
    public:
        VectorMD() = default;

        template <typename... Args, typename = std::enable_if_t<sizeof...(Args) == N>>
        explicit VectorMD(Args... dims)
        {
            resize(dims...);
        }

        template <typename U, typename... Args, typename = std::enable_if_t<sizeof...(Args) == N>>
        VectorMD(const Fill<U> init, Args... dims)
        {
            resize(dims...);
            std::fill(data_.begin(), data_.end(), static_cast<T>(init.val_));
        }

        template <typename... Args>
        void resize(Args... dims)
        {
            static_assert(sizeof...(Args) == N, "one extent per dimension");
            std::size_t i = 0;
            ((dims_[i++] = static_cast<std::size_t>(dims)), ...);
            std::size_t stride = 1;
            for (std::size_t k = N; k-- > 0;)
            {
                strides_[k] = stride;
                stride *= dims_[k];
            }
            data_.assign(stride, T{});
        }

        template <typename... Idx>
        T &operator()(Idx... idx)
        {
            return data_[offset(idx...)];
        }

        template <typename... Idx>
        const T &operator()(Idx... idx) const
        {
            return data_[offset(idx...)];
        }

        T &operator[](std::size_t i) { return data_[i]; }
        const T &operator[](std::size_t i) const { return data_[i]; }

        template <std::size_t k>
        std::size_t dim() const
        {
            static_assert(k < N, "dimension out of range");
            return dims_[k];
        }

        std::size_t size() const { return data_.size(); }
        T *memptr() { return data_.data(); }
        const T *memptr() const { return data_.data(); }
        typename std::vector<T>::iterator begin() { return data_.begin(); }
        typename std::vector<T>::iterator end() { return data_.end(); }
        typename std::vector<T>::const_iterator begin() const { return data_.begin(); }
        typename std::vector<T>::const_iterator end() const { return data_.end(); }

    private:
        template <typename... Idx>
        std::size_t offset(Idx... idx) const
        {
            static_assert(sizeof...(Idx) == N, "one index per dimension");
            std::size_t o = 0, k = 0;
            ((o += static_cast<std::size_t>(idx) * strides_[k++]), ...);
            return o;
        }

        std::array<std::size_t, N> dims_{};
        std::array<std::size_t, N> strides_{};
        std::vector<T> data_;
    };

    using vec2d = VectorMD<double, 2>;
    using vec3d = VectorMD<double, 3>;
    using vec4d = VectorMD<double, 4>;
}
