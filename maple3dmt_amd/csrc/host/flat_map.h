// Open-addressing hash map for the mesh enumeration: keys to non-negative
// int ids, linear probing in a power-of-two table kept at most half full.
// Entries live in two flat arrays, so an insert allocates nothing (the
// std::unordered_map it replaces allocated one node per entry, millions per
// mesh).  Ids are whatever the caller stores, so first-appearance numbering
// does not depend on the table.
#pragma once

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

namespace m3d {

inline uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

struct Hash64 {
    uint64_t operator()(uint64_t k) const { return mix64(k); }
};

template <typename Key, typename Hash = Hash64>
class FlatMap {
public:
    explicit FlatMap(size_t expected = 16) { rehash(capacity_for(expected)); }

    // The id stored for k, or -1.
    int find(const Key& k) const {
        size_t i = Hash()(k) & mask_;
        while (val_[i] >= 0) {
            if (key_[i] == k) return val_[i];
            i = (i + 1) & mask_;
        }
        return -1;
    }

    // {id of k, false} when present, else stores v and returns {v, true}.
    std::pair<int, bool> insert(const Key& k, int v) {
        if (2 * (size_ + 1) > val_.size()) rehash(val_.size() * 2);
        size_t i = Hash()(k) & mask_;
        while (val_[i] >= 0) {
            if (key_[i] == k) return {val_[i], false};
            i = (i + 1) & mask_;
        }
        key_[i] = k;
        val_[i] = v;
        ++size_;
        return {v, true};
    }

    size_t size() const { return size_; }

private:
    static size_t capacity_for(size_t n) {
        size_t c = 16;
        while (c < 2 * n) c <<= 1;
        return c;
    }
    void rehash(size_t cap) {
        std::vector<Key> ok;
        std::vector<int> ov;
        ok.swap(key_);
        ov.swap(val_);
        key_.assign(cap, Key{});
        val_.assign(cap, -1);
        mask_ = cap - 1;
        size_ = 0;
        for (size_t i = 0; i < ov.size(); ++i)
            if (ov[i] >= 0) insert(ok[i], ov[i]);
    }
    std::vector<Key> key_;
    std::vector<int> val_;
    size_t mask_ = 0, size_ = 0;
};

}  // namespace m3d
