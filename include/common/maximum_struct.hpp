// Peak record of the kept host API (common/maximum_struct.hpp:3-6): grid
// indices in the consumer's Float type and the field value as float.
#pragma once

template <typename T> struct Maximum {
    T i0, i1, i2;
    float field;
};
