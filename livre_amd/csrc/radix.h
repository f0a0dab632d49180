// Stable LSD radix sort of (key, value) pairs by an unsigned integer key,
// 11-bit digits, as many passes as the largest key needs. Used where the
// host orders a few thousand bricks per frame (selection output, draw
// order); the callers fall back to std::sort wherever equal keys would make
// the result depend on how ties are broken.
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

namespace lvr
{
template <typename K, typename V>
void radixSortByKey(std::vector<std::pair<K, V>>& a, std::vector<std::pair<K, V>>& tmp)
{
    K maxKey = 0;
    for (const auto& p : a)
        maxKey = p.first > maxKey ? p.first : maxKey;
    tmp.resize(a.size());
    constexpr int kBits = 11;
    constexpr uint32_t kMask = (1u << kBits) - 1u;
    for (int shift = 0; shift < int(8 * sizeof(K)) && (maxKey >> shift) != 0; shift += kBits)
    {
        uint32_t count[kMask + 2] = {};
        for (const auto& p : a)
            ++count[((p.first >> shift) & kMask) + 1];
        for (uint32_t d = 0; d <= kMask; ++d)
            count[d + 1] += count[d];
        for (const auto& p : a)
            tmp[count[(p.first >> shift) & kMask]++] = p;
        a.swap(tmp);
    }
}

// true when two neighbours of a key-sorted sequence share a key
template <typename K, typename V>
bool hasEqualKeys(const std::vector<std::pair<K, V>>& a)
{
    for (size_t i = 1; i < a.size(); ++i)
        if (a[i].first == a[i - 1].first)
            return true;
    return false;
}
} // namespace lvr
