// Drop-in NeighborSet (reference inc/structures/neighbor_set.hpp:15-172): half-open index ranges
// into the caller's sorted cloud, forward iteration over (index, point).
#pragma once

#include <cstddef>
#include <cstdint>
#include <iterator>
#include <type_traits>
#include <utility>
#include <vector>

#include "../geometry/point_containers.hpp"

template <class Container>
class NeighborSet {
public:
    Container *points = nullptr;
    std::vector<std::pair<size_t, size_t>> ranges;
    size_t numberOfPoints = 0;

    NeighborSet() = default;
    explicit NeighborSet(Container *pts) : points(pts) {}

    void addRange(size_t first, size_t last) {
        if (first <= last) {
            ranges.emplace_back(first, last);
            numberOfPoints += last - first;
        }
    }
    size_t size() const { return numberOfPoints; }
    bool empty() const { return numberOfPoints == 0; }

    // (index, point): the point by reference for containers that store Points (PointsAoS), by value
    // for those that assemble them on access (PointsSoA)
    using point_ref = std::conditional_t<std::is_reference_v<decltype(std::declval<const Container &>()[0])>, const Point &, Point>;

    class Iterator {
    public:
        using iterator_category = std::forward_iterator_tag;
        using value_type = std::pair<size_t, point_ref>;
        using difference_type = std::ptrdiff_t;
        using pointer = value_type *;
        using reference = value_type;

        Iterator(const NeighborSet &s, size_t r) : s_(s), r_(r), i_(r < s.ranges.size() ? s.ranges[r].first : SIZE_MAX) {
            settle();
        }
        reference operator*() const { return {i_, (*s_.points)[i_]}; }
        Iterator &operator++() {
            ++i_;
            settle();
            return *this;
        }
        Iterator operator++(int) {
            Iterator t = *this;
            ++(*this);
            return t;
        }
        bool operator==(const Iterator &o) const { return i_ == o.i_; }
        bool operator!=(const Iterator &o) const { return i_ != o.i_; }

    private:
        const NeighborSet &s_;
        size_t r_, i_;
        void settle() {  // skip exhausted (and empty) ranges
            while (r_ < s_.ranges.size() && i_ >= s_.ranges[r_].second) {
                if (++r_ < s_.ranges.size()) i_ = s_.ranges[r_].first;
            }
            if (r_ >= s_.ranges.size()) i_ = SIZE_MAX;
        }
    };

    Iterator begin() const { return Iterator(*this, 0); }
    Iterator end() const { return Iterator(*this, ranges.size()); }
};
