// CtVolume-compatible C++ facade over libctg (the engine API that Ct's GUI and CLI call,
// Ct/CtVolume.h:57-278), without Qt or OpenCV: std::string for QString, callbacks for
// the Qt signals, images decoded by imageio.h.  Method names, argument meaning,
// defaults and error reporting follow the reference; reconstruction runs on the GPU
// through the C ABI in include/ctg.h.
#pragma once

#include <atomic>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "imageio.h"

struct ctg_context;

namespace ct {

// An allocator whose value-initialisation leaves trivial elements uninitialised:
// resize() of a multi-GB volume then costs no single-threaded zero fill (the facade
// zeroes the volume on worker threads instead).
template <class T>
struct NoInitAllocator : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInitAllocator<U>;
    };
    NoInitAllocator() = default;
    template <class U>
    NoInitAllocator(const NoInitAllocator<U>&) noexcept {}
    template <class U>
    void construct(U* p) noexcept {
        ::new (static_cast<void*>(p)) U;
    }
    template <class U, class... A>
    void construct(U* p, A&&... a) {
        ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
    }
};
using VolumeData = std::vector<float, NoInitAllocator<float>>;

// Ct/CtVolume.h:50-54
enum class FilterType { RAMLAK, SHEPP_LOGAN, HANN };
// Common/Volume.h:18-32
enum class Axis { X, Y, Z };
enum class IndexOrder { X_FASTEST, Z_FASTEST };
// QDataStream::ByteOrder values
enum class ByteOrder { BigEndian = 0, LittleEndian = 1 };

// Common/types.h:11-20
struct CompletionStatus {
    bool successful = true;
    bool userInterrupted = false;
    std::string errorMessage;
    static CompletionStatus success() { return {true, false, {}}; }
    static CompletionStatus interrupted() { return {false, true, {}}; }
    static CompletionStatus error(const std::string& m) { return {false, false, m}; }
};

// Ct/CtVolume.h:42-48 (the image is the raw decoded projection)
struct Projection {
    Image image;
    double angle = 0.0;
    double heightOffset = 0.0;
};

// The raw stream of Volume::saveToBinaryFile (Common/Volume.h:474-574) for a Z_FASTEST
// volume in memory: 32-bit floats in the given index and byte order.  Returns 0 on
// success, 1 when the file cannot be opened, 2 on a write error; progress (may be empty)
// receives percentages, and a set *stop (may be null) ends the write early.
int writeRawVolume(const std::string& path, const float* zfastest, size_t nx, size_t ny, size_t nz,
                   IndexOrder order, ByteOrder byteOrder, const std::function<void(double)>& progress,
                   const std::atomic<bool>* stop);

class CtVolume {
public:
    using ProgressFn = std::function<void(double percentage)>;
    // reconstructionProgress(double, cv::Mat crossSection) (Ct/CtVolume.h:273): the
    // cross-section at the set axis and index of the volume as far as it is finished
    using CrossSectionProgressFn = std::function<void(double percentage, const Image& crossSection)>;
    using FinishedFn = std::function<void(const CompletionStatus&)>;

    CtVolume();
    explicit CtVolume(const std::string& csvFile);
    ~CtVolume();
    CtVolume(const CtVolume&) = delete;
    CtVolume& operator=(const CtVolume&) = delete;

    // getters (Ct/CtVolume.h:67-94)
    static bool cudaAvailable(bool verbose = false);
    static int getCudaDeviceCount(bool verbose = false);
    Projection getProjectionAt(size_t index) const;
    size_t getSinogramSize() const;
    unsigned getImageWidth() const;
    unsigned getImageHeight() const;
    unsigned getXSize() const;
    unsigned getYSize() const;
    unsigned getZSize() const;
    unsigned getReconstructionCylinderRadius() const;
    float getUOffset() const;
    float getPixelSize() const;
    float getFCD() const;
    float getBaseIntensity() const;
    bool getEmitSignals() const;
    bool getUseCuda() const;
    unsigned getSkipProjections() const;
    std::vector<int> getActiveCudaDevices() const;
    std::vector<std::string> getCudaDeviceList() const;
    size_t getGpuSpareMemory() const;
    double getMultiprocessorCoefficient() const;
    double getMemoryBandwidthCoefficient() const;
    size_t getRequiredMemoryUpperBound() const;
    bool getUseCpuPreprocessing() const;
    unsigned getCrossSectionIndex() const;
    unsigned getCrossSectionSize() const;
    Axis getCrossSectionAxis() const;
    // A slice of the reconstructed volume as a 32-bit float image in the left-handed
    // orientation (Ct/CtVolume.cpp:391-393, Common/Volume.h:577-641): Axis::X gives
    // rows = z (top row = highest z) by columns = y; Axis::Y gives rows = z by columns =
    // x, highest x first; Axis::Z gives rows = x by columns = y.  Empty before a
    // reconstruction; throws std::out_of_range for an index outside the axis.
    Image getVolumeCrossSection(Axis axis, unsigned index) const;

    // setters (Ct/CtVolume.h:97-113)
    void setCrossSectionIndex(unsigned index);
    void setCrossSectionAxis(Axis axis);
    void setEmitSignals(bool value);
    void setVolumeBounds(float xFrom, float xTo, float yFrom, float yTo, float zFrom, float zTo);
    void setUseCuda(bool value);
    void setActiveCudaDevices(std::vector<int> devices);
    void setGpuSpareMemory(size_t memoryMb);
    void setGpuCoefficients(double multiprocessorCoefficient, double memoryBandwidthCoefficient);
    void setUseCpuPreprocessing(bool value);
    void setFrequencyFilterType(FilterType filterType);
    void setSkipProjections(unsigned value = 0);

    // control (Ct/CtVolume.h:116-121)
    bool sinogramFromImages(const std::string& csvFile);
    void reconstructVolume();
    void saveVolumeToBinaryFile(const std::string& filename, IndexOrder indexOrder = IndexOrder::Z_FASTEST,
                                ByteOrder byteOrder = ByteOrder::LittleEndian) const;
    void stop();

    // stand-ins for the Qt signals (Ct/CtVolume.h:270-277)
    void onLoadingProgress(ProgressFn f) { loadingProgress_ = std::move(f); }
    void onLoadingFinished(FinishedFn f) { loadingFinished_ = std::move(f); }
    void onReconstructionProgress(ProgressFn f) { reconstructionProgress_ = std::move(f); }
    // With a cross-section: called after every finished z-slab (or x range) of the
    // reconstruction, as the reference's CUDA path emits one after every chunk
    // (Ct/CtVolume.cpp:1470-1485).  Voxels of slabs not yet finished are 0.
    void onReconstructionProgress(CrossSectionProgressFn f) { reconstructionCrossSection_ = std::move(f); }
    void onReconstructionFinished(FinishedFn f) { reconstructionFinished_ = std::move(f); }
    void onSavingProgress(ProgressFn f) { savingProgress_ = std::move(f); }
    void onSavingFinished(FinishedFn f) { savingFinished_ = std::move(f); }

    // the reconstructed volume, Z_FASTEST [x][y][z] (empty before a reconstruction)
    const VolumeData& volumeData() const { return volume_; }
    std::string lastErrorMessage() const { return lastErrorMessage_; }
    CompletionStatus lastStatus() const { return lastStatus_; }

private:
    struct Entry {
        std::string imagePath;
        double angle;
        double heightOffset;
    };
    void makeHeightOffsetRelative();
    void correctAngleDirection(const std::string& rotationDirection);
    void updateBoundaries();
    void finishLoading(const CompletionStatus& s);
    void finishReconstruction(const CompletionStatus& s);
    void finishSaving(const CompletionStatus& s) const;

    std::vector<Entry> sinogram_;
    unsigned xSize_ = 0, ySize_ = 0, zSize_ = 0;
    unsigned imageWidth_ = 0, imageHeight_ = 0;
    PixelDepth imageDepth_ = PixelDepth::U16;
    bool mixedDepths_ = false;
    float FCD_ = 0, pixelSize_ = 0, uOffset_ = 0, baseIntensity_ = 0;

    FilterType filterType_ = FilterType::RAMLAK;
    unsigned projectionStep_ = 1;
    bool emitSignals_ = true;
    unsigned crossSectionIndex_ = 0;
    Axis crossSectionAxis_ = Axis::Z;
    bool useCuda_ = true;
    std::vector<int> activeCudaDevices_{0};
    size_t gpuSpareMemory_ = 200;
    double multiprocessorCoefficient_ = 1, memoryBandwidthCoefficient_ = 1;
    bool useCpuPreprocessing_ = false;
    mutable std::atomic<bool> stopActiveProcess_{false};
    float xFrom_float_ = 0, xTo_float_ = 1, yFrom_float_ = 0, yTo_float_ = 1, zFrom_float_ = 0, zTo_float_ = 1;

    VolumeData volume_;
    // The images the last successful load decoded, in sinogram order and native byte
    // order, kept when they share one depth and fit a third of the host's available
    // memory (imageCacheBytes_ == 0 otherwise, or with CTG_FACADE_IMAGE_CACHE=0): the
    // reconstruction then takes them instead of decoding every file a second time.
    std::unique_ptr<uint8_t[]> imageCache_;
    size_t imageCacheBytes_ = 0;
    unsigned xFrom_ = 0, xTo_ = 0, yFrom_ = 0, yTo_ = 0, zFrom_ = 0, zTo_ = 0;
    unsigned xMax_ = 0, yMax_ = 0, zMax_ = 0;
    std::pair<float, float> minMaxValues_{0.0f, 0.0f};
    mutable std::mutex exclusiveFunctionsMutex_;
    std::string lastErrorMessage_;
    mutable CompletionStatus lastStatus_;
    // the context of a running reconstruction; stop() cancels it under contextMutex_,
    // and reconstructVolume clears it under the same lock before destroying it
    std::mutex contextMutex_;
    ctg_context* activeContext_ = nullptr;

    ProgressFn loadingProgress_, reconstructionProgress_;
    CrossSectionProgressFn reconstructionCrossSection_;
    mutable ProgressFn savingProgress_;
    FinishedFn loadingFinished_, reconstructionFinished_;
    mutable FinishedFn savingFinished_;
};

}  // namespace ct
