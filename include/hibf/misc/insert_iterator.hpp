// seqan::hibf::insert_iterator for the MI355X-native HIBF: the output iterator
// a config::input_fn writes user-bin values into (misc/insert_iterator.hpp:23-98
// of the reference).  Here it appends to a host vector (or calls a function);
// the build then hands all values to the device bulk emplace.
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <iterator>
#include <memory>
#include <vector>

namespace seqan::hibf
{

class insert_iterator
{
public:
    using iterator_category = std::output_iterator_tag;
    using value_type = void;
    using difference_type = std::ptrdiff_t;
    using pointer = void;
    using reference = void;
    using function_t = std::function<void(uint64_t const)>;

    insert_iterator() = default;
    explicit insert_iterator(std::vector<uint64_t> & sink) : sink_{std::addressof(sink)} {}
    explicit insert_iterator(function_t & fun) : fun_{std::addressof(fun)} {}

    insert_iterator & operator=(uint64_t const value)
    {
        if (sink_)
            sink_->push_back(value);
        else if (fun_)
            (*fun_)(value);
        return *this;
    }
    insert_iterator & operator*() noexcept { return *this; }
    insert_iterator & operator++() noexcept { return *this; }
    insert_iterator operator++(int) noexcept { return *this; }

private:
    std::vector<uint64_t> * sink_{nullptr};
    function_t * fun_{nullptr};
};

} // namespace seqan::hibf
