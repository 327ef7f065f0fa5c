// mgpusort.hpp -- C++ wrapper over the mgpusort C API.
//
// Same classes and member functions as the reference's inc/mgpusort.hpp
// (MgpuException :12-20, MgpuSortData :25-75, MgpuSort :79-114) so that C++
// clients compile unchanged.
#ifndef MGPUSORT_HPP
#define MGPUSORT_HPP

#include <cstring>
#include <exception>
#include <utility>

#include "mgpusort.h"

// Thrown by the wrappers for any status other than SORT_STATUS_SUCCESS.
struct MgpuException : std::exception {
    explicit MgpuException(sortStatus_t status) : _status(status) {}
    const char* what() const throw() override {
        const char* s = sortStatusString(_status);
        return s ? s : "unknown sortStatus_t";
    }
    sortStatus_t _status;
};

// sortData_d that remembers which buffers the caller attached, so that Reset
// (and the destructor) frees only what sortAllocData allocated.  Attached
// buffers are tracked by their position at attach time; since sorts swap the
// pairs, position i at attach time is position i ^ parity now.
struct MgpuSortData : sortData_d {
    MgpuSortData() { clear_all(); }
    ~MgpuSortData() { Reset(); }

    MgpuSortData(const MgpuSortData&) = delete;
    MgpuSortData& operator=(const MgpuSortData&) = delete;

    void Reset() {
        for (int i = 0; i < 2; ++i) {
            if (_attachedKeys[i]) keys[i ^ parity] = 0;
            for (int j = 0; j < 6; ++j)
                if (_attachedVals[j][i]) values[j][i ^ parity] = 0;
        }
        sortFreeData(this);
        clear_all();
    }

    sortStatus_t Alloc(sortEngine_t engine, int maxElements_, int valueCount_) {
        maxElements = maxElements_;
        numElements = maxElements_;
        valueCount = valueCount_;
        return sortAllocData(engine, this);
    }

    void AttachKey(CUdeviceptr p, int i = 0) {
        keys[i] = p;
        _attachedKeys[i ^ parity] = true;
    }
    void AttachKey(unsigned int* p, int i = 0) {
        pKeys[i] = p;
        _attachedKeys[i ^ parity] = true;
    }
    void AttachVal(int j, CUdeviceptr p, int i = 0) {
        values[j][i] = p;
        _attachedVals[j][i ^ parity] = true;
    }
    void AttachVal(int j, unsigned int* p, int i = 0) {
        pValues[j][i] = p;
        _attachedVals[j][i ^ parity] = true;
    }

private:
    void clear_all() {
        std::memset(static_cast<sortData_d*>(this), 0, sizeof(sortData_d));
        std::memset(_attachedKeys, 0, sizeof(_attachedKeys));
        std::memset(_attachedVals, 0, sizeof(_attachedVals));
    }
    bool _attachedKeys[2];
    bool _attachedVals[6][2];
};

// Reference-counted engine handle: copies share the engine.
class MgpuSort {
public:
    explicit MgpuSort(const char* kernelPath) {
        sortStatus_t st = sortCreateEngine(kernelPath, &_engine);
        if (st != SORT_STATUS_SUCCESS) throw MgpuException(st);
    }
    // Wraps an existing handle and takes a reference to it.
    MgpuSort(sortEngine_t engine) : _engine(engine) {
        if (_engine) sortIncEngine(_engine);
    }
    MgpuSort(const MgpuSort& rhs) : _engine(rhs._engine) {
        if (_engine) sortIncEngine(_engine);
    }
    ~MgpuSort() {
        if (_engine) sortReleaseEngine(_engine);
    }
    MgpuSort& operator=(const MgpuSort& rhs) {
        MgpuSort tmp(rhs);
        std::swap(_engine, tmp._engine);
        return *this;
    }
    operator sortEngine_t() { return _engine; }

private:
    sortEngine_t _engine = nullptr;
};

#endif  // MGPUSORT_HPP
