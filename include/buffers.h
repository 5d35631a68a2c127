/*
 * buffers.h -- the reference's Vector type (src/buffers.h:9-50), so that
 * callers written against PartConvFilter::setup(M, N, num_parts, Vector*, ...)
 * (src/partconv.h:48-53) and readImpulseResponse(Vector&, ...)
 * (src/sndtools.h:16) compile unchanged.
 *
 * VectorTemplate<T> is a zero-initialised, owning array of `size` elements with
 * the reference's member functions: operator[] (element), operator()(i)
 * (pointer to element i), create, destroy, reset, getSize, zeroPad, copyFrom.
 * Host memory only (the engine copies an impulse response to the device at
 * setup).  The reference's FFT buffer types (ComplexVector, the Matrix
 * templates) belong to its FFTW engine and have no counterpart here.
 */
#ifndef PARTCONV_AMD_BUFFERS_H
#define PARTCONV_AMD_BUFFERS_H

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

template <class T>
class VectorTemplate
{
public:
    VectorTemplate(int inSize = 0) : size(0), data(nullptr) { create(inSize); }
    VectorTemplate(const VectorTemplate<T> &vec) : size(0), data(nullptr)
    {
        create(vec.size);
        if (size > 0) memcpy(data, vec.data, sizeof(T) * size);
    }
    ~VectorTemplate() { destroy(); }
    VectorTemplate<T> &operator=(const VectorTemplate<T> &vec)
    {
        if (this != &vec) {
            create(vec.size);
            if (size > 0) memcpy(data, vec.data, sizeof(T) * size);
        }
        return *this;
    }
    T &operator[](int ind) { return data[ind]; }
    const T &operator[](int ind) const { return data[ind]; }
    T *operator()(int ind = 0) { return data + ind; }
    const T *operator()(int ind = 0) const { return data + ind; }

    /* (Re)size to inSize zeroed elements: 0 on success, 1 for inSize <= 0
     * (the vector is then empty) or an allocation failure. */
    int create(int inSize)
    {
        if (inSize == size && inSize > 0) return 0;
        destroy();
        if (inSize <= 0) return 1;
        data = static_cast<T *>(calloc((size_t)inSize, sizeof(T)));
        if (!data) return 1;
        size = inSize;
        return 0;
    }
    void destroy()
    {
        free(data);
        data = nullptr;
        size = 0;
    }
    void reset()
    {
        if (size > 0) memset(data, 0, sizeof(T) * size);
    }
    int getSize() const { return size; }

    /* Grow to `length` elements with the old contents at `offset` and zeros
     * elsewhere; 1 (unchanged) if length < size + offset. */
    int zeroPad(int length, int offset = 0)
    {
        if (length < size + offset || offset < 0) {
            fprintf(stderr, "VectorTemplate<T>::zeroPad() padded size too small!\n");
            return 1;
        }
        T *nd = static_cast<T *>(calloc((size_t)(length > 0 ? length : 1), sizeof(T)));
        if (!nd) return 1;
        if (size > 0) memcpy(nd + offset, data, sizeof(T) * size);
        free(data);
        data = nd;
        size = length;
        return 0;
    }
    void copyFrom(const T *src)
    {
        if (size > 0) memcpy(data, src, sizeof(T) * size);
    }

private:
    int size;
    T *data;
};

typedef VectorTemplate<float> Vector;

#endif /* PARTCONV_AMD_BUFFERS_H */
