// arena.hpp -- placement of job buffers inside one reserved block of device
// memory (host-side bookkeeping only; the block itself is a device pointer).
#pragma once

#include <algorithm>
#include <cstddef>
#include <iterator>
#include <map>

namespace rtm_hip {

// A reserved block of device memory with first-fit placement and coalescing
// on release (offset-ordered free list); blocks are 256-byte aligned.
class Arena {
public:
    bool active() const { return base_ != nullptr; }
    size_t size() const { return size_; }
    bool contains(const void* p) const {
        return base_ && static_cast<const char*>(p) >= base_ && static_cast<const char*>(p) < base_ + size_;
    }
    void reset(char* base, size_t size) {
        base_ = base;
        size_ = size;
        free_.clear();
        used_.clear();
        if (base) free_[0] = size;
    }
    char* base() const { return base_; }
    void* alloc(size_t n) {
        n = (std::max<size_t>(n, 1) + 255) & ~size_t(255);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            const size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > n) free_[off + n] = len - n;
            used_[off] = n;
            return base_ + off;
        }
        return nullptr;
    }
    void release(const void* p) {
        const size_t off = (size_t)(static_cast<const char*>(p) - base_);
        auto u = used_.find(off);
        if (u == used_.end()) return;
        size_t len = u->second;
        size_t start = off;
        used_.erase(u);
        auto next = free_.lower_bound(off);
        if (next != free_.end() && next->first == off + len) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                start = prev->first;
                len += prev->second;
                free_.erase(prev);
            }
        }
        free_[start] = len;
    }

private:
    char* base_ = nullptr;
    size_t size_ = 0;
    std::map<size_t, size_t> free_;   // offset -> length
    std::map<size_t, size_t> used_;
};

}  // namespace rtm_hip
