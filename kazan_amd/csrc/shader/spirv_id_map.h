// spirv_id_map.h — a table indexed by SPIR-V id, bounded by the module
// header's `bound`.
//
// Counterpart of the reference's IdMap (spirv-id-map/src/lib.rs:94-210): a
// dense array of optional values at index id - 1, sized from the bound, so
// that every id a module uses must lie in [1, bound) (ids outside it are
// "out of bounds", an error the translator reports). The interface is the
// subset of std::map the translator uses (find / count / at / operator[] /
// erase / iteration in id order); operator[] on an id outside the bound
// returns a scratch slot and marks the map, which the translator checks.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <utility>
#include <vector>

namespace kz {

template <class V>
class IdMap {
  public:
    using value_type = std::pair<uint32_t, V>;

    IdMap() = default;
    explicit IdMap(uint32_t bound) { reset(bound); }

    // empties the map and sets its bound (ids 1 .. bound - 1)
    void reset(uint32_t bound) {
        slots_.assign(bound > 0 ? bound - 1 : 0, value_type{});
        used_.assign(slots_.size(), 0);
        for (size_t i = 0; i < slots_.size(); ++i) slots_[i].first = (uint32_t)i + 1;
        out_of_bounds_ = 0;
    }

    bool in_bounds(uint32_t id) const { return id >= 1 && (size_t)id - 1 < slots_.size(); }
    // an id used outside the bound, 0 if none
    uint32_t out_of_bounds() const { return out_of_bounds_; }

    class iterator {
      public:
        iterator(IdMap *m, size_t i) : m_(m), i_(i) { skip(); }
        value_type &operator*() const { return m_->slots_[i_]; }
        value_type *operator->() const { return &m_->slots_[i_]; }
        iterator &operator++() {
            ++i_;
            skip();
            return *this;
        }
        bool operator==(const iterator &o) const { return i_ == o.i_; }
        bool operator!=(const iterator &o) const { return i_ != o.i_; }

      private:
        friend class IdMap;
        void skip() {
            while (i_ < m_->slots_.size() && !m_->used_[i_]) ++i_;
        }
        IdMap *m_;
        size_t i_;
    };
    class const_iterator {
      public:
        const_iterator(const IdMap *m, size_t i) : m_(m), i_(i) { skip(); }
        const value_type &operator*() const { return m_->slots_[i_]; }
        const value_type *operator->() const { return &m_->slots_[i_]; }
        const_iterator &operator++() {
            ++i_;
            skip();
            return *this;
        }
        bool operator==(const const_iterator &o) const { return i_ == o.i_; }
        bool operator!=(const const_iterator &o) const { return i_ != o.i_; }

      private:
        void skip() {
            while (i_ < m_->slots_.size() && !m_->used_[i_]) ++i_;
        }
        const IdMap *m_;
        size_t i_;
    };

    iterator begin() { return iterator(this, 0); }
    iterator end() { return iterator(this, slots_.size()); }
    const_iterator begin() const { return const_iterator(this, 0); }
    const_iterator end() const { return const_iterator(this, slots_.size()); }

    iterator find(uint32_t id) { return in_bounds(id) && used_[id - 1] ? iterator(this, id - 1) : end(); }
    const_iterator find(uint32_t id) const {
        return in_bounds(id) && used_[id - 1] ? const_iterator(this, id - 1) : end();
    }
    size_t count(uint32_t id) const { return in_bounds(id) && used_[id - 1] ? 1 : 0; }

    V &at(uint32_t id) {
        if (!count(id)) throw std::out_of_range("SPIR-V id not defined");
        return slots_[id - 1].second;
    }
    const V &at(uint32_t id) const {
        if (!count(id)) throw std::out_of_range("SPIR-V id not defined");
        return slots_[id - 1].second;
    }

    V &operator[](uint32_t id) {
        if (!in_bounds(id)) {
            if (!out_of_bounds_) out_of_bounds_ = id ? id : ~0u;
            scratch_ = V{};
            return scratch_;
        }
        used_[id - 1] = 1;
        return slots_[id - 1].second;
    }

    void erase(uint32_t id) {
        if (count(id)) {
            used_[id - 1] = 0;
            slots_[id - 1].second = V{};
        }
    }
    iterator erase(iterator it) {
        erase(it->first);
        ++it;
        return it;
    }

  private:
    std::vector<value_type> slots_;
    std::vector<uint8_t> used_;
    V scratch_{};
    uint32_t out_of_bounds_ = 0;
};

}  // namespace kz
