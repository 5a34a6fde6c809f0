// CtVolume-compatible facade over libctg; see ctvolume.h.  Behaviour restated from
// Ct/CtVolume.cpp (file:line cited per method) and Common/Volume.h (writer).
#include "ctvolume.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <thread>

#include <unistd.h>

#include "../../include/ctg.h"

namespace ct {

namespace {

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

// QString::toDouble: whitespace-trimmed, the whole string must parse.
double to_double(const std::string& s, bool& ok) {
    const std::string t = trim(s);
    if (t.empty()) {
        ok = false;
        return 0.0;
    }
    char* end = nullptr;
    const double v = std::strtod(t.c_str(), &end);
    ok = end && *end == '\0';
    return ok ? v : 0.0;
}

std::string section0(const std::string& line) { return line.substr(0, line.find('\t')); }

std::vector<std::string> split_tabs(const std::string& line) {
    std::vector<std::string> out;
    size_t start = 0;
    while (true) {
        size_t p = line.find('\t', start);
        out.push_back(line.substr(start, p == std::string::npos ? std::string::npos : p - start));
        if (p == std::string::npos) break;
        start = p + 1;
    }
    return out;
}

bool is_absolute(const std::string& p) { return !p.empty() && p[0] == '/'; }

std::string dir_of(const std::string& path) {
    const size_t p = path.find_last_of('/');
    if (p == std::string::npos) return ".";
    return p == 0 ? "/" : path.substr(0, p);
}

std::string file_of(const std::string& path) {
    const size_t p = path.find_last_of('/');
    return p == std::string::npos ? path : path.substr(p + 1);
}

std::string join(const std::string& dir, const std::string& name) {
    if (is_absolute(name)) return name;
    if (dir.empty() || dir == ".") return name;
    return dir.back() == '/' ? dir + name : dir + "/" + name;
}

// QDir::cleanPath of an absolute path
std::string clean_path(const std::string& path) {
    std::vector<std::string> parts;
    std::stringstream ss(path);
    std::string part;
    while (std::getline(ss, part, '/')) {
        if (part.empty() || part == ".") continue;
        if (part == "..") {
            if (!parts.empty()) parts.pop_back();
        } else {
            parts.push_back(part);
        }
    }
    std::string out;
    for (auto& p : parts) out += "/" + p;
    return out.empty() ? "/" : out;
}

std::string absolute(const std::string& path) {
    if (is_absolute(path)) return clean_path(path);
    char buf[4096];
    const char* cwd = getcwd(buf, sizeof(buf));
    return clean_path(std::string(cwd ? cwd : ".") + "/" + path);
}

// printf("%g"), the formatting of QTextStream << float and QString::arg(double)
std::string fmt_g(double v) {
    char b[64];
    std::snprintf(b, sizeof(b), "%g", v);
    return b;
}

std::string fmt_f3(double v) {
    char b[64];
    std::snprintf(b, sizeof(b), "%.3f", v);
    return b;
}

double pixel_scale(PixelDepth d) { return d == PixelDepth::U8 ? 255.0 : (d == PixelDepth::U16 ? 65535.0 : 1.0); }

// Decodes paths[i] into imgs[i] with a pool of threads; on_image(i, image), when given,
// runs on the decoding thread after each successful decode.
void decode_all(const std::vector<std::string>& paths, std::vector<Image>& imgs, std::vector<std::string>& errs,
                const std::function<void(size_t, const Image&)>& on_image = nullptr) {
    const size_t n = paths.size();
    imgs.assign(n, Image());
    errs.assign(n, std::string());
    const unsigned nt = std::max(1u, std::min<unsigned>(16, std::thread::hardware_concurrency()));
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; ++t) {
        pool.emplace_back([&]() {
            for (size_t i = next++; i < n; i = next++) {
                if (!readImage(paths[i], imgs[i], errs[i])) imgs[i] = Image();
                else if (on_image) on_image(i, imgs[i]);
            }
        });
    }
    for (auto& th : pool) th.join();
}

// Runs fn(begin, end) on parts of [0, n) of at least min_part elements, on up to 16
// threads (the calling thread takes the first part).
template <class F>
void parallel_parts(size_t n, size_t min_part, F&& fn) {
    const size_t hw = std::max(1u, std::min<unsigned>(16, std::thread::hardware_concurrency()));
    const size_t parts = std::max<size_t>(1, std::min<size_t>(hw, n / std::max<size_t>(1, min_part)));
    const size_t step = (n + parts - 1) / parts;
    std::vector<std::thread> pool;
    for (size_t t = 1; t < parts; ++t) {
        const size_t b = t * step, e = std::min(n, b + step);
        if (b < e) pool.emplace_back([&fn, b, e]() { fn(b, e); });
    }
    fn(0, std::min(n, step));
    for (auto& th : pool) th.join();
}

// MemAvailable of /proc/meminfo in bytes (0 when it cannot be read).
size_t host_memory_available() {
    std::ifstream f("/proc/meminfo");
    std::string line;
    while (std::getline(f, line)) {
        unsigned long long kb = 0;
        if (std::sscanf(line.c_str(), "MemAvailable: %llu kB", &kb) == 1) return (size_t)kb * 1024;
    }
    return 0;
}

// the load keeps its decoded images for the reconstruction unless CTG_FACADE_IMAGE_CACHE=0
bool image_cache_enabled() {
    const char* e = std::getenv("CTG_FACADE_IMAGE_CACHE");
    return !(e && std::string(e) == "0");
}

// Per-image minimum and maximum as Ct/CtVolume.cpp:271-281 scale them (by 256 for 8-bit
// and 65536 for 16-bit and float images).
std::pair<double, double> image_min_max(const Image& im) {
    double lMin = std::numeric_limits<double>::infinity(), lMax = -lMin;
    const size_t npx = (size_t)im.width * im.height;
    for (size_t i = 0; i < npx; ++i) {
        const double v = im.depth == PixelDepth::U8    ? im.data[i]
                         : im.depth == PixelDepth::U16 ? reinterpret_cast<const uint16_t*>(im.data.data())[i]
                                                       : reinterpret_cast<const float*>(im.data.data())[i];
        lMin = std::min(lMin, v);
        lMax = std::max(lMax, v);
    }
    const double div = im.depth == PixelDepth::U8 ? 256.0 : 65536.0;
    return {lMin / div, lMax / div};
}

}  // namespace

// ------------------------------------------------------------------- getters

CtVolume::CtVolume() = default;

CtVolume::CtVolume(const std::string& csvFile) { sinogramFromImages(csvFile); }

CtVolume::~CtVolume() = default;

bool CtVolume::cudaAvailable(bool verbose) { return getCudaDeviceCount(verbose) > 0; }

int CtVolume::getCudaDeviceCount(bool verbose) {
    int32_t n = 0;
    if (ctg_device_count(&n) != CTG_OK || n <= 0) {
        if (verbose) std::cout << "No HIP devices could be found. GPU reconstruction will not be available." << std::endl;
        return 0;
    }
    return n;
}

Projection CtVolume::getProjectionAt(size_t index) const {
    if (index >= sinogram_.size()) throw std::out_of_range("Index out of bounds.");
    Projection p;
    std::string err;
    readImage(sinogram_[index].imagePath, p.image, err);
    p.angle = sinogram_[index].angle;
    p.heightOffset = sinogram_[index].heightOffset;
    return p;
}

size_t CtVolume::getSinogramSize() const { return sinogram_.size(); }
unsigned CtVolume::getImageWidth() const { return imageWidth_; }
unsigned CtVolume::getImageHeight() const { return imageHeight_; }
unsigned CtVolume::getXSize() const { return xMax_; }
unsigned CtVolume::getYSize() const { return yMax_; }
unsigned CtVolume::getZSize() const { return zMax_; }

// Ct/CtVolume.cpp:370-373
unsigned CtVolume::getReconstructionCylinderRadius() const {
    const double radius = imageWidth_ / 2;
    return (unsigned)std::sqrt(((double)(FCD_ * FCD_) * radius * radius) / ((double)(FCD_ * FCD_) + radius * radius));
}

float CtVolume::getUOffset() const { return uOffset_; }
float CtVolume::getPixelSize() const { return pixelSize_; }
float CtVolume::getFCD() const { return FCD_; }
float CtVolume::getBaseIntensity() const { return baseIntensity_; }
bool CtVolume::getEmitSignals() const { return emitSignals_; }
bool CtVolume::getUseCuda() const { return useCuda_; }
unsigned CtVolume::getSkipProjections() const { return projectionStep_ - 1; }
std::vector<int> CtVolume::getActiveCudaDevices() const { return activeCudaDevices_; }

std::vector<std::string> CtVolume::getCudaDeviceList() const {
    const int n = getCudaDeviceCount();
    std::vector<std::string> out;
    for (int i = 0; i < n; ++i) {
        ctg_device_info info;
        if (ctg_get_device_info(i, &info) == CTG_OK) {
            std::ostringstream s;
            s.setf(std::ios::fixed);
            s.precision(1);
            s << info.name << ", " << (double)info.total_bytes / 1024.0 / 1024.0 / 1024.0 << " Gb";
            out.push_back(s.str());
        } else {
            out.push_back("unavailable");
        }
    }
    return out;
}

size_t CtVolume::getGpuSpareMemory() const { return gpuSpareMemory_; }
double CtVolume::getMultiprocessorCoefficient() const { return multiprocessorCoefficient_; }
double CtVolume::getMemoryBandwidthCoefficient() const { return memoryBandwidthCoefficient_; }

// Ct/CtVolume.cpp:469-481
size_t CtVolume::getRequiredMemoryUpperBound() const {
    size_t required = (size_t)xMax_ * yMax_ * zMax_ * sizeof(float);
    if (useCuda_) required += (size_t)imageWidth_ * imageHeight_ * sizeof(float) * activeCudaDevices_.size() * 3;
    else required += (size_t)imageWidth_ * imageHeight_ * sizeof(float) * 2;
    return required;
}

bool CtVolume::getUseCpuPreprocessing() const { return useCpuPreprocessing_; }
unsigned CtVolume::getCrossSectionIndex() const { return crossSectionIndex_; }

unsigned CtVolume::getCrossSectionSize() const {
    return crossSectionAxis_ == Axis::X ? xMax_ : (crossSectionAxis_ == Axis::Y ? yMax_ : zMax_);
}

Axis CtVolume::getCrossSectionAxis() const { return crossSectionAxis_; }

// Ct/CtVolume.cpp:391-393 -> Common/Volume.h:577-641 with LEFT_HANDED.  volume_ is
// Z_FASTEST, element (x, y, z) at (x * yMax + y) * zMax + z.
Image CtVolume::getVolumeCrossSection(Axis axis, unsigned index) const {
    Image out;
    if (volume_.empty() || xMax_ == 0) return out;
    const unsigned n = axis == Axis::X ? xMax_ : (axis == Axis::Y ? yMax_ : zMax_);
    if (index >= n) throw std::out_of_range("Index out of bounds.");
    const size_t ny = yMax_, nz = zMax_;
    const int rows = axis == Axis::Z ? (int)xMax_ : (int)zMax_;
    const int cols = axis == Axis::Y ? (int)xMax_ : (int)yMax_;
    out.width = cols;
    out.height = rows;
    out.depth = PixelDepth::F32;
    out.data.resize((size_t)rows * cols * sizeof(float));
    float* dst = reinterpret_cast<float*>(out.data.data());
    for (int r = 0; r < rows; ++r) {
        float* row = dst + (size_t)r * cols;
        if (axis == Axis::Z) {
            // contiguous in y at stride nz: (x = r, y = c, z = index)
            const float* src = volume_.data() + (size_t)r * ny * nz + index;
            for (int c = 0; c < cols; ++c) row[c] = src[(size_t)c * nz];
        } else if (axis == Axis::X) {
            // (x = index, y = c, z = rows - 1 - r)
            const float* src = volume_.data() + (size_t)index * ny * nz + (size_t)(rows - 1 - r);
            for (int c = 0; c < cols; ++c) row[c] = src[(size_t)c * nz];
        } else {
            // (x = cols - 1 - c, y = index, z = rows - 1 - r)
            const float* src = volume_.data() + (size_t)index * nz + (size_t)(rows - 1 - r);
            for (int c = 0; c < cols; ++c) row[c] = src[(size_t)(cols - 1 - c) * ny * nz];
        }
    }
    return out;
}

// ------------------------------------------------------------------- setters

void CtVolume::setCrossSectionIndex(unsigned index) {
    if ((crossSectionAxis_ == Axis::X && index < xMax_) || (crossSectionAxis_ == Axis::Y && index < yMax_) ||
        (crossSectionAxis_ == Axis::Z && index < zMax_))
        crossSectionIndex_ = index;
}

void CtVolume::setCrossSectionAxis(Axis axis) {
    crossSectionAxis_ = axis;
    crossSectionIndex_ = (axis == Axis::X ? xMax_ : (axis == Axis::Y ? yMax_ : zMax_)) / 2;
}

void CtVolume::setEmitSignals(bool value) { emitSignals_ = value; }

// Ct/CtVolume.cpp:487-500 (yTo and zTo are clamped against xFrom, as in the reference)
void CtVolume::setVolumeBounds(float xFrom, float xTo, float yFrom, float yTo, float zFrom, float zTo) {
    std::lock_guard<std::mutex> lock(exclusiveFunctionsMutex_);
    if (xFrom == xTo || yFrom == yTo || zFrom == zTo) {
        std::cout << "ERROR: the lower and upper bounds must not be identical." << std::endl;
        return;
    }
    xFrom_float_ = std::max(0.0f, std::min(1.0f, xFrom));
    xTo_float_ = std::max(xFrom_float_, std::min(1.0f, xTo));
    yFrom_float_ = std::max(0.0f, std::min(1.0f, yFrom));
    yTo_float_ = std::max(xFrom_float_, std::min(1.0f, yTo));
    zFrom_float_ = std::max(0.0f, std::min(1.0f, zFrom));
    zTo_float_ = std::max(xFrom_float_, std::min(1.0f, zTo));
    if (!sinogram_.empty()) updateBoundaries();
}

void CtVolume::setUseCuda(bool value) { useCuda_ = value; }

// Ct/CtVolume.cpp:506-520
void CtVolume::setActiveCudaDevices(std::vector<int> devices) {
    const int n = getCudaDeviceCount();
    devices.erase(std::remove_if(devices.begin(), devices.end(), [n](int d) { return d >= n; }), devices.end());
    if (!devices.empty()) activeCudaDevices_ = devices;
    else std::cout << "Active CUDA devices were not set because vector did not contain any valid device ID." << std::endl;
}

void CtVolume::setGpuSpareMemory(size_t memoryMb) { gpuSpareMemory_ = memoryMb; }

void CtVolume::setGpuCoefficients(double mp, double bw) {
    multiprocessorCoefficient_ = mp;
    memoryBandwidthCoefficient_ = bw;
}

void CtVolume::setUseCpuPreprocessing(bool value) { useCpuPreprocessing_ = value; }
void CtVolume::setFrequencyFilterType(FilterType filterType) { filterType_ = filterType; }
void CtVolume::setSkipProjections(unsigned value) { projectionStep_ = value + 1; }

void CtVolume::stop() {
    stopActiveProcess_ = true;
    std::lock_guard<std::mutex> lock(contextMutex_);
    if (activeContext_) ctg_cancel(activeContext_);
}

// ---------------------------------------------------------------- internals

void CtVolume::finishLoading(const CompletionStatus& s) {
    lastStatus_ = s;
    if (emitSignals_ && loadingFinished_) loadingFinished_(s);
}

void CtVolume::finishReconstruction(const CompletionStatus& s) {
    lastStatus_ = s;
    if (emitSignals_ && reconstructionFinished_) reconstructionFinished_(s);
}

void CtVolume::finishSaving(const CompletionStatus& s) const {
    lastStatus_ = s;
    if (emitSignals_ && savingFinished_) savingFinished_(s);
}

// Ct/CtVolume.cpp:742-753
void CtVolume::makeHeightOffsetRelative() {
    double sum = 0;
    for (auto& e : sinogram_) sum += e.heightOffset;
    sum /= (double)sinogram_.size();
    for (auto& e : sinogram_) {
        e.heightOffset -= sum;
        e.heightOffset /= pixelSize_;
    }
}

// Ct/CtVolume.cpp:755-767
void CtVolume::correctAngleDirection(const std::string& rotationDirection) {
    if (sinogram_.size() > 1) {
        const double diff = sinogram_[1].angle - sinogram_[0].angle;
        if ((rotationDirection == "cw" && diff > 0) || (rotationDirection == "ccw" && diff < 0))
            for (auto& e : sinogram_) e.angle *= -1;
    }
}

// Ct/CtVolume.cpp:1399-1424
void CtVolume::updateBoundaries() {
    xFrom_ = (unsigned)((double)(xFrom_float_ * (float)xSize_) + 0.5);
    xTo_ = (unsigned)((double)(xTo_float_ * (float)xSize_) + 0.5);
    yFrom_ = (unsigned)((double)(yFrom_float_ * (float)ySize_) + 0.5);
    yTo_ = (unsigned)((double)(yTo_float_ * (float)ySize_) + 0.5);
    zFrom_ = (unsigned)((double)(zFrom_float_ * (float)zSize_) + 0.5);
    zTo_ = (unsigned)((double)(zTo_float_ * (float)zSize_) + 0.5);
    xMax_ = xTo_ - xFrom_;
    yMax_ = yTo_ - yFrom_;
    zMax_ = zTo_ - zFrom_;
    if (crossSectionAxis_ == Axis::X && crossSectionIndex_ >= xMax_) crossSectionIndex_ = xMax_ / 2;
    else if (crossSectionAxis_ == Axis::Y && crossSectionIndex_ >= yMax_) crossSectionIndex_ = yMax_ / 2;
    else if (crossSectionAxis_ == Axis::Z && crossSectionIndex_ >= zMax_) crossSectionIndex_ = zMax_ / 2;
}

// ----------------------------------------------------------------- loading

// Ct/CtVolume.cpp:136-330
bool CtVolume::sinogramFromImages(const std::string& csvFile) {
    std::lock_guard<std::mutex> lock(exclusiveFunctionsMutex_);
    stopActiveProcess_ = false;
    volume_.clear();
    sinogram_.clear();
    imageCache_.reset();
    imageCacheBytes_ = 0;
    std::ifstream in(csvFile);
    if (!in) {
        std::cerr << "Could not open CSV file - terminating" << std::endl;
        finishLoading(CompletionStatus::error("Could not open the config file."));
        return false;
    }
    auto read_line = [&](std::string& line) -> bool {
        if (!std::getline(in, line)) {
            line.clear();
            return false;
        }
        if (!line.empty() && line.back() == '\r') line.pop_back();
        return true;
    };
    std::string line, rotationDirection, imageDir;
    {
        bool ok, total = true;
        read_line(line);
        imageDir = section0(line);
        read_line(line);
        pixelSize_ = (float)to_double(section0(line), ok);
        total = total && ok;
        read_line(line);
        rotationDirection = section0(line);
        read_line(line);
        uOffset_ = (float)to_double(section0(line), ok);
        total = total && ok;
        read_line(line);
        FCD_ = (float)to_double(section0(line), ok);
        total = total && ok;
        read_line(line);
        baseIntensity_ = (float)to_double(section0(line), ok);
        total = total && ok;
        read_line(line);  // one ignored line
        FCD_ /= pixelSize_;
        uOffset_ /= pixelSize_;
        if (!total) {
            const std::string m = "Could not read the parameters from the CSV file successfully.";
            std::cout << m << std::endl;
            finishLoading(CompletionStatus::error(m));
            return false;
        } else if (baseIntensity_ == 0.0f) {
            const std::string m = "Error: The base intensity given in the configuration file must not be equal to zero.";
            std::cout << m << std::endl;
            finishLoading(CompletionStatus::error(m));
            return false;
        }
    }
    if (!is_absolute(imageDir)) imageDir = clean_path(join(dir_of(absolute(csvFile)), imageDir));
    {
        bool ok, total = true, have = false;
        std::string imageFilename;
        double angle = 0.0, heightOffset = 0.0;
        while (read_line(line)) {
            const std::vector<std::string> fields = split_tabs(line);
            if (fields.size() >= 2) {
                imageFilename = fields[0];
                angle = to_double(fields[1], ok);
                total = total && ok;
                have = true;
            } else if (!have) {
                // the reference would push an entry with unset name and angle here
                total = false;
            }
            if (fields.size() >= 3) {
                heightOffset = to_double(fields[2], ok);
                total = total && ok;
            } else {
                heightOffset = 0;
            }
            // a line with fewer than two fields repeats the previous name and angle
            sinogram_.push_back({join(imageDir, imageFilename), angle, heightOffset});
        }
        if (!total) {
            sinogram_.clear();
            const std::string m = "Could not read the image parameters from the CSV file successfully.";
            std::cout << m << std::endl;
            finishLoading(CompletionStatus::error(m));
            return false;
        }
        if (sinogram_.empty()) {
            const std::string m = "Apparently the config file does not contain any images.";
            std::cout << m << std::endl;
            finishLoading(CompletionStatus::error(m));
            return false;
        }
    }
    // load and check images (decoded in parallel, checked in order)
    {
        std::vector<std::string> paths;
        for (auto& e : sinogram_) paths.push_back(e.imagePath);
        std::vector<Image> imgs;
        std::vector<std::string> errs;
        const size_t chunk = 64;
        unsigned rows = 0, cols = 0;
        double mn = std::numeric_limits<double>::quiet_NaN(), mx = std::numeric_limits<double>::quiet_NaN();
        mixedDepths_ = false;
        // The decoded images are kept for the reconstruction when they all share the
        // first image's size and depth and fit the budget (decided from the first image).
        std::unique_ptr<uint8_t[]> cache;
        size_t cacheStride = 0;
        PixelDepth cacheDepth = PixelDepth::U16;
        if (image_cache_enabled() && !paths.empty()) {
            Image first;
            std::string e;
            if (readImage(paths[0], first, e) && !first.empty()) {
                const size_t bytes = first.data.size() * paths.size();
                if (bytes <= host_memory_available() / 3) {
                    cache.reset(new (std::nothrow) uint8_t[bytes]);
                    if (cache) {
                        cacheStride = first.data.size();
                        cacheDepth = first.depth;
                    }
                }
            }
        }
        std::atomic<bool> cacheOk{cache != nullptr};
        std::vector<std::pair<double, double>> minMax(chunk);
        for (size_t c0 = 0; c0 < paths.size(); c0 += chunk) {
            if (stopActiveProcess_) {
                sinogram_.clear();
                std::cout << "User interrupted. Stopping.";
                finishLoading(CompletionStatus::interrupted());
                return false;
            }
            std::vector<std::string> part(paths.begin() + c0, paths.begin() + std::min(paths.size(), c0 + chunk));
            // on the decoding threads: the image's min/max, and its copy into the cache
            decode_all(part, imgs, errs, [&](size_t k, const Image& im) {
                minMax[k] = image_min_max(im);
                if (cacheOk.load(std::memory_order_relaxed)) {
                    if (im.depth == cacheDepth && im.data.size() == cacheStride)
                        std::memcpy(cache.get() + (c0 + k) * cacheStride, im.data.data(), cacheStride);
                    else
                        cacheOk = false;
                }
            });
            for (size_t k = 0; k < part.size(); ++k) {
                const size_t cnt = c0 + k;
                const Image& im = imgs[k];
                if (im.empty()) {
                    const std::string m = "Error loading the image \"" + part[k] +
                                          "\". Maybe it does not exist, permissions are missing or the format is "
                                          "not supported.";
                    sinogram_.clear();
                    std::cout << m << std::endl;
                    finishLoading(CompletionStatus::error(m));
                    return false;
                }
                if (cnt == 0) {
                    rows = (unsigned)im.height;
                    cols = (unsigned)im.width;
                    imageDepth_ = im.depth;
                } else if ((unsigned)im.height != rows || (unsigned)im.width != cols) {
                    const std::string m =
                        "Error loading the image \"" + part[k] + "\", its dimensions differ from the images before.";
                    sinogram_.clear();
                    std::cout << m << std::endl;
                    finishLoading(CompletionStatus::error(m));
                    return false;
                }
                if (im.depth != imageDepth_) mixedDepths_ = true;
                // per-image min/max scaled as in Ct/CtVolume.cpp:271-281 (computed on the
                // decoding thread), combined in image order
                const double lMin = minMax[k].first, lMax = minMax[k].second;
                if (std::isnan(mn) || lMin < mn) mn = lMin;
                if (std::isnan(mx) || lMax > mx) mx = lMax;
                imageWidth_ = cols;
                imageHeight_ = rows;
                const double pct = std::round(double(cnt) / double(sinogram_.size()) * 100);
                std::cout << "\r" << "Analysing images: " << pct << "%";
                if (emitSignals_ && loadingProgress_) loadingProgress_(pct);
            }
        }
        minMaxValues_ = {(float)mn, (float)mx};
        if ((float)imageWidth_ >= FCD_) {
            sinogram_.clear();
            const std::string m = "Error: the FCD must be greater than the image width.";
            std::cout << m << std::endl;
            finishLoading(CompletionStatus::error(m));
            return false;
        }
        std::cout << std::endl;
        if (cacheOk && !mixedDepths_) {
            imageCache_ = std::move(cache);
            imageCacheBytes_ = cacheStride * paths.size();
        }
    }
    makeHeightOffsetRelative();
    correctAngleDirection(rotationDirection);
    const unsigned radius = getReconstructionCylinderRadius();
    xSize_ = radius * 2;
    ySize_ = radius * 2;
    zSize_ = imageHeight_;
    updateBoundaries();
    crossSectionIndex_ = (crossSectionAxis_ == Axis::X ? xMax_ : (crossSectionAxis_ == Axis::Y ? yMax_ : zMax_)) / 2;
    finishLoading(CompletionStatus::success());
    return true;
}

// ----------------------------------------------------------- reconstruction

// Ct/CtVolume.cpp:543-632 with the GPU engine replaced by ctg_reconstruct.
void CtVolume::reconstructVolume() {
    std::lock_guard<std::mutex> lock(exclusiveFunctionsMutex_);
    lastErrorMessage_ = "An error during the reconstruction occured.";
    stopActiveProcess_ = false;
    if (sinogram_.empty()) {
        const std::string m =
            "Volume was not reconstructed, because the sinogram seems to be empty. Please load some images first.";
        std::cout << m << std::endl;
        finishReconstruction(CompletionStatus::error(m));
        return;
    }
    volume_.clear();
    try {
        volume_.resize((size_t)xMax_ * yMax_ * zMax_);
    } catch (...) {
        const std::string m = "The memory allocation for the volume failed. Maybe there is not enought free RAM.";
        std::cout << m << std::endl;
        finishReconstruction(CompletionStatus::error(m));
        return;
    }
    // zeroed on worker threads (the slabs not yet reconstructed read as 0 in the live
    // cross-section); the first touch of the pages is spread over the threads too
    {
        float* v = volume_.data();
        parallel_parts(volume_.size(), (size_t)1 << 22, [v](size_t b, size_t e) { std::fill(v + b, v + e, 0.0f); });
    }
    const auto start = std::chrono::steady_clock::now();
    auto fail = [&](const std::string& m) {
        volume_.clear();
        lastErrorMessage_ = m;
        std::cout << std::endl << m << std::endl;
        finishReconstruction(CompletionStatus::error(m));
    };
    if (!useCuda_) {
        fail("Reconstruction runs on the GPU in ct_amd; the CPU path (useCuda = false, -n) is not provided.");
        return;
    }
    if (!cudaAvailable()) {
        fail("GPU processing was requested, but no capable HIP device could be found.");
        return;
    }
    if (activeCudaDevices_.size() > 1)
        std::cout << "Using device " << activeCudaDevices_[0] << " (one device per reconstruction in ct_amd)."
                  << std::endl;

    // projections actually used (projectionStep, Ct/CtVolume.cpp:945)
    std::vector<size_t> used;
    for (size_t p = 0; p < sinogram_.size(); p += projectionStep_) used.push_back(p);
    const size_t npx = (size_t)imageWidth_ * imageHeight_;
    const PixelDepth depth = mixedDepths_ ? PixelDepth::F32 : imageDepth_;
    const size_t bpp = depth == PixelDepth::U8 ? 1 : (depth == PixelDepth::U16 ? 2 : 4);
    // The images: from the load's cache when it holds this sinogram (every image, one
    // depth), gathered when only every projectionStep-th is used; otherwise decoded here.
    std::vector<uint8_t, NoInitAllocator<uint8_t>> raw;
    const uint8_t* images = nullptr;
    const bool cached = imageCacheBytes_ > 0 && !mixedDepths_ && imageCacheBytes_ == sinogram_.size() * npx * bpp;
    if (!cached || used.size() != sinogram_.size()) {
        try {
            raw.resize(used.size() * npx * bpp);
        } catch (...) {
            fail("The memory allocation for the projections failed. Maybe there is not enought free RAM.");
            return;
        }
    }
    if (cached && used.size() == sinogram_.size()) {
        images = imageCache_.get();
    } else if (cached) {
        const uint8_t* src = imageCache_.get();
        uint8_t* dst = raw.data();
        const size_t ib = npx * bpp;
        parallel_parts(used.size(), 1, [&](size_t b, size_t e) {
            for (size_t k = b; k < e; ++k) std::memcpy(dst + k * ib, src + used[k] * ib, ib);
        });
        images = raw.data();
    } else {
        const size_t chunk = 64;
        std::vector<Image> imgs;
        std::vector<std::string> errs;
        for (size_t c0 = 0; c0 < used.size(); c0 += chunk) {
            if (stopActiveProcess_) {
                volume_.clear();
                std::cout << std::endl << "User interrupted. Stopping." << std::endl;
                finishReconstruction(CompletionStatus::interrupted());
                return;
            }
            std::vector<std::string> part;
            for (size_t k = c0; k < std::min(used.size(), c0 + chunk); ++k) part.push_back(sinogram_[used[k]].imagePath);
            decode_all(part, imgs, errs);
            for (size_t k = 0; k < part.size(); ++k) {
                const Image& im = imgs[k];
                if (im.empty() || (unsigned)im.width != imageWidth_ || (unsigned)im.height != imageHeight_) {
                    fail("The image " + part[k] +
                         " could not be accessed. Maybe it doesn't exist or has an unsupported format.");
                    return;
                }
                uint8_t* dst = raw.data() + (c0 + k) * npx * bpp;
                if (!mixedDepths_) {
                    memcpy(dst, im.data.data(), npx * bpp);
                } else {
                    // Mixed depths: intensities as fractions of full scale, 32-bit float.
                    // For the CPU log form an integer image also takes the 1e-6*B that
                    // convertTo32bit adds to 8U and 16U images only (Ct/CtVolume.cpp:
                    // 802-812), evaluated as the GPU does for an integer image,
                    // fma(I, 1/S, 1e-6*B) in float, so a zero pixel stays finite and the
                    // values equal those of an unmixed integer stack.
                    float* d = reinterpret_cast<float*>(dst);
                    const double s = pixel_scale(im.depth);
                    const bool integer = im.depth != PixelDepth::F32;
                    const float la = (float)(1.0 / s);
                    const float lb = (float)(0.000001 * (double)baseIntensity_);
                    for (size_t i = 0; i < npx; ++i) {
                        const double v = im.depth == PixelDepth::U8 ? im.data[i]
                                         : im.depth == PixelDepth::U16
                                             ? reinterpret_cast<const uint16_t*>(im.data.data())[i]
                                             : reinterpret_cast<const float*>(im.data.data())[i];
                        if (integer && useCpuPreprocessing_) d[i] = std::fma((float)v, la, lb);
                        else d[i] = (float)(v / s);
                    }
                }
            }
        }
        images = raw.data();
    }
    std::vector<double> angles, heights;
    for (size_t p : used) {
        angles.push_back(sinogram_[p].angle);
        heights.push_back(sinogram_[p].heightOffset);
    }
    ctg_context* ctx = nullptr;
    if (ctg_create(activeCudaDevices_.empty() ? 0 : activeCudaDevices_[0], &ctx) != CTG_OK) {
        fail(std::string("The GPU could not be initialised: ") + ctg_last_error());
        return;
    }
    {
        std::lock_guard<std::mutex> lock(contextMutex_);
        activeContext_ = ctx;
        if (stopActiveProcess_) ctg_cancel(ctx);
    }
    const float bounds[6] = {xFrom_float_, xTo_float_, yFrom_float_, yTo_float_, zFrom_float_, zTo_float_};
    ctg_geometry g;
    uint32_t roi[6];
    ctg_status st = ctg_geometry_ctvolume(imageWidth_, imageHeight_, FCD_, uOffset_, bounds, &g, roi);
    if (st == CTG_OK) st = ctg_set_geometry(ctx, &g);
    if (st == CTG_OK) st = ctg_set_projections(ctx, angles.data(), heights.data(), (int32_t)used.size());
    if (st == CTG_OK) st = ctg_set_spare_memory(ctx, gpuSpareMemory_);
    if (st == CTG_OK) {
        struct Cb {
            CtVolume* self;
        } cb{this};
        auto progress = [](double f, void* user) {
            CtVolume* self = static_cast<Cb*>(user)->self;
            const double pct = std::round(f * 100.0);
            std::cout << "\r" << "Total completion: " << pct << "%" << std::flush;
            if (self->emitSignals_ && self->reconstructionProgress_) self->reconstructionProgress_(pct);
            // the slabs finished so far are in the host volume (their downloads have
            // returned before this call); the rest of it is still zero
            if (self->emitSignals_ && self->reconstructionCrossSection_)
                self->reconstructionCrossSection_(pct, self->getVolumeCrossSection(self->crossSectionAxis_,
                                                                                   self->crossSectionIndex_));
        };
        const ctg_pixel_type pix =
            depth == PixelDepth::U8 ? CTG_PIXEL_U8 : (depth == PixelDepth::U16 ? CTG_PIXEL_U16 : CTG_PIXEL_F32);
        const ctg_filter filt = filterType_ == FilterType::RAMLAK        ? CTG_FILTER_RAMLAK
                                : filterType_ == FilterType::SHEPP_LOGAN ? CTG_FILTER_SHEPP_LOGAN
                                                                          : CTG_FILTER_HANN;
        // preprocessing always runs on the GPU; the CPU-preprocessing switch selects the
        // CPU path's log form (Ct/CtVolume.cpp:802-812 vs 892-904)
        st = ctg_reconstruct(ctx, images, pix, (int32_t)used.size(), baseIntensity_,
                             useCpuPreprocessing_ ? CTG_LOG_CPU_FORM : CTG_LOG_GPU_FORM, filt, CTG_STORE_F32,
                             volume_.data(), CTG_Z_FASTEST, progress, &cb);
    }
    const std::string err = ctg_last_error();
    {
        std::lock_guard<std::mutex> lock(contextMutex_);
        activeContext_ = nullptr;
    }
    ctg_destroy(ctx);
    if (st == CTG_ERROR_INTERRUPTED || stopActiveProcess_) {
        volume_.clear();
        std::cout << std::endl << "User interrupted. Stopping." << std::endl;
        finishReconstruction(CompletionStatus::interrupted());
        return;
    }
    if (st != CTG_OK) {
        fail(err.empty() ? lastErrorMessage_ : err);
        return;
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    std::cout << std::endl << "Volume successfully reconstructed (" << secs << "s)" << std::endl;
    finishReconstruction(CompletionStatus::success());
}

// ------------------------------------------------------------------ writer

// The raw stream of Volume::saveToBinaryFile (Common/Volume.h:474-574): every element
// as a 32-bit IEEE float in the requested byte order, z fastest (element (x, y, z) at
// (x*ny + y)*nz + z, the in-memory order) or x fastest ((z*ny + y)*nx + x).  The
// reference writes element by element through QDataStream; here the stream is cut into
// chunks of about 32 MiB that worker threads fill (a cache-blocked transpose for X
// fastest, a byte swap for big endian) while a writer thread writes the previous chunk,
// so the file write and the reordering overlap.  Returns 0 on success, 1 when the file
// cannot be opened, 2 on a write error; a set *stop ends the write early (returning 0).
int writeRawVolume(const std::string& path, const float* zfastest, size_t nx, size_t ny, size_t nz,
                   IndexOrder order, ByteOrder byteOrder, const std::function<void(double)>& progress,
                   const std::atomic<bool>* stop) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return 1;
    const bool swap = byteOrder == ByteOrder::BigEndian;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(zfastest);
    // a chunk is a range of whole x planes (Z fastest) or whole z slices (X fastest)
    const size_t plane = order == IndexOrder::Z_FASTEST ? ny * nz : nx * ny;
    const size_t nplanes = order == IndexOrder::Z_FASTEST ? nx : nz;
    size_t per_chunk = std::max<size_t>(1, ((size_t)32 << 20) / 4 / std::max<size_t>(1, plane));
    // X fastest reads 16 consecutive z of a column per chunk (one 64-byte line)
    if (order == IndexOrder::X_FASTEST) per_chunk = std::max<size_t>(16, per_chunk / 16 * 16);
    const unsigned nt = std::max(1u, std::min<unsigned>(16, std::thread::hardware_concurrency()));
    std::vector<uint32_t> buf[2];
    bool ok = true;
    std::thread writer;
    auto fill = [&](std::vector<uint32_t>& out, size_t p0, size_t np) {
        out.resize(np * plane);
        // work split into units: x planes (Z fastest) or 8-row y bands of the slices (X fastest)
        const size_t yb = 8;
        const size_t nby = (ny + yb - 1) / yb;
        const size_t units = order == IndexOrder::Z_FASTEST ? np : nby;
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (size_t u = next++; u < units; u = next++) {
                if (order == IndexOrder::Z_FASTEST) {
                    const uint32_t* s = src + (p0 + u) * plane;
                    uint32_t* d = out.data() + u * plane;
                    if (swap)
                        for (size_t i = 0; i < plane; ++i) d[i] = __builtin_bswap32(s[i]);
                    else
                        std::memcpy(d, s, plane * 4);
                } else {
                    // slices z in [p0, p0+np), rows y in [y0, y1): tiles of 16 x by the run of z
                    const size_t y0 = u * yb, y1 = std::min(ny, y0 + yb);
                    for (size_t x0 = 0; x0 < nx; x0 += 16) {
                        const size_t x1 = std::min(nx, x0 + 16);
                        for (size_t y = y0; y < y1; ++y)
                            for (size_t x = x0; x < x1; ++x) {
                                const uint32_t* col = src + (x * ny + y) * nz + p0;
                                uint32_t* d = out.data() + y * nx + x;
                                for (size_t k = 0; k < np; ++k) {
                                    const uint32_t w = col[k];
                                    d[k * plane] = swap ? __builtin_bswap32(w) : w;
                                }
                            }
                    }
                }
            }
        };
        const unsigned nw = (unsigned)std::min<size_t>(nt, units);
        std::vector<std::thread> pool;
        for (unsigned t = 1; t < nw; ++t) pool.emplace_back(work);
        work();
        for (auto& t : pool) t.join();
    };
    size_t k = 0;
    for (size_t p0 = 0; p0 < nplanes; p0 += per_chunk, ++k) {
        if (stop && stop->load()) break;
        const size_t np = std::min(per_chunk, nplanes - p0);
        fill(buf[k & 1], p0, np);
        if (writer.joinable()) writer.join();  // the previous chunk is on disk; its buffer is free
        if (!ok) break;
        writer = std::thread([&, k]() {
            const std::vector<uint32_t>& b = buf[k & 1];
            if (std::fwrite(b.data(), 4, b.size(), f) != b.size()) ok = false;
        });
        if (progress) progress(std::round(double(p0) / double(nplanes) * 100));
    }
    if (writer.joinable()) writer.join();
    const bool closed = std::fclose(f) == 0;
    return ok && closed ? 0 : 2;
}

// Ct/CtVolume.cpp:634-721 and Common/Volume.h:474-574
void CtVolume::saveVolumeToBinaryFile(const std::string& filename, IndexOrder indexOrder, ByteOrder byteOrder) const {
    std::lock_guard<std::mutex> lock(exclusiveFunctionsMutex_);
    stopActiveProcess_ = false;
    const std::string dir = dir_of(filename), fname = file_of(filename);
    const std::string baseName = fname.substr(0, fname.find('.'));
    {
        std::ofstream out(join(dir, baseName + ".txt"), std::ios::binary | std::ios::trunc);
        if (!out) {
            std::cout << "Could not write the info file." << std::endl;
            finishSaving(CompletionStatus::error("Could not write the info file."));
            return;
        }
        out << fname << "\n\n";
        out << "[Image dimensions]\n";
        out << "U resolution:\t\t" << imageWidth_ << "\n";
        out << "V resolution:\t\t" << imageHeight_ << "\n\n";
        out << "[Reconstruction parameters]\n";
        out << "FCD:\t\t\t\t\t" << fmt_g(FCD_) << "\n";
        out << "Pixel size:\t\t\t" << fmt_g(pixelSize_) << "\n";
        out << "U offset:\t\t\t" << fmt_g(uOffset_) << "\n";
        out << "X range relative:\t[" << fmt_f3(xFrom_float_) << " - " << fmt_f3(xTo_float_) << "]\n";
        out << "Y range relative:\t[" << fmt_f3(yFrom_float_) << " - " << fmt_f3(yTo_float_) << "]\n";
        out << "Z range relative:\t[" << fmt_f3(zFrom_float_) << " - " << fmt_f3(zTo_float_) << "]\n";
        out << "X range:\t\t\t[" << xFrom_ << ".." << xTo_ << "]\n";
        out << "Y range:\t\t\t[" << yFrom_ << ".." << yTo_ << "]\n";
        out << "Z range:\t\t\t[" << zFrom_ << ".." << zTo_ << "]\n\n";
        out << "[Volume dimensions]\n";
        out << "X size:\t\t\t\t" << xMax_ << "\n";
        out << "Y size:\t\t\t\t" << yMax_ << "\n";
        out << "Z size:\t\t\t\t" << zMax_ << "\n\n";
        out << "[Data format]\n";
        out << "Data type:\t\t\t32bit IEEE 754 float\n";
        out << (byteOrder == ByteOrder::LittleEndian ? "Byte order:\t\t\tLittle endian\n" : "Byte order:\t\t\tBig endian\n");
        out << (indexOrder == IndexOrder::Z_FASTEST ? "Index order:\t\tZ fastest" : "Index order:\t\tX fastest");
    }
    float vmin = 0.0f, vmax = 0.0f;
    if (!volume_.empty()) {
        // the volume's range for the .vgi, over parts of it on worker threads
        const float* v = volume_.data();
        std::mutex mm_lock;
        vmin = vmax = v[0];
        parallel_parts(volume_.size(), (size_t)1 << 22, [&](size_t b, size_t e) {
            auto mm = std::minmax_element(v + b, v + e);
            std::lock_guard<std::mutex> g(mm_lock);
            vmin = std::min(vmin, *mm.first);
            vmax = std::max(vmax, *mm.second);
        });
    }
    {
        std::ofstream out(join(dir, baseName + ".vgi"), std::ios::binary | std::ios::trunc);
        if (!out) {
            std::cout << "Could not write the info file." << std::endl;
            finishSaving(CompletionStatus::error("Could not write the vgi file."));
            return;
        }
        const std::string X = std::to_string(xMax_), Y = std::to_string(yMax_), Z = std::to_string(zMax_);
        const std::string size = X + " " + Y + " " + Z, range = fmt_g(vmin) + " " + fmt_g(vmax);
        out << "{volume1}\n[representation]\nsize = " << size << "\nbitsperelement = 32\ndatatype = float\n"
            << "datarange = " << range << "\n[file1]\nFileFormat = raw\nSize = " << size << "\nName = ./" << fname
            << "\nbitsperelement = 32\ndatatype = float\ndatarange = " << range << "\n{volumeprimitive}\n[geometry]\n"
            << "clipbox = 0 0 0 " << size << "\nstatus = visible\nresolution = " << fmt_g(pixelSize_) << " "
            << fmt_g(pixelSize_) << " " << fmt_g(pixelSize_) << "\nunit = mm\n[volume]\nvolume = volume1\n";
    }
    // the raw stream (Common/Volume.h:474-574)
    if (volume_.empty() || xMax_ == 0 || yMax_ == 0 || zMax_ == 0) {
        std::cout << "Did not save the volume, because it appears to be empty." << std::endl;
        finishSaving(CompletionStatus::error("Did not save the volume, because it appears to be empty."));
        return;
    }
    auto progress = [this](double pct) {
        if (emitSignals_ && savingProgress_) savingProgress_(pct);
    };
    const int rc = writeRawVolume(filename, volume_.data(), xMax_, yMax_, zMax_, indexOrder, byteOrder, progress,
                                  &stopActiveProcess_);
    if (rc == 1) {
        const std::string m = "Could not open the file. Maybe your path does not exist. No files were written.";
        std::cout << m << std::endl;
        finishSaving(CompletionStatus::error(m));
        return;
    }
    const bool ok = rc == 0;
    if (stopActiveProcess_) {
        std::cout << "User interrupted. Stopping." << std::endl;
        finishSaving(CompletionStatus::interrupted());
        return;
    }
    if (!ok) {
        const std::string m = "An error occured while writing to the disk. Maybe there is not enough free disk space.";
        std::cout << m << std::endl;
        finishSaving(CompletionStatus::error(m));
        return;
    }
    std::cout << "Volume successfully saved." << std::endl;
    finishSaving(CompletionStatus::success());
}

}  // namespace ct

// --------------------------------------------------------- C wrapper (tests)

extern "C" {

void* ctv_create(void) { return new ct::CtVolume(); }
void ctv_destroy(void* v) { delete static_cast<ct::CtVolume*>(v); }
int ctv_load(void* v, const char* csv) { return static_cast<ct::CtVolume*>(v)->sinogramFromImages(csv) ? 1 : 0; }
void ctv_set_bounds(void* v, const float* b) {
    static_cast<ct::CtVolume*>(v)->setVolumeBounds(b[0], b[1], b[2], b[3], b[4], b[5]);
}
void ctv_set_filter(void* v, int f) { static_cast<ct::CtVolume*>(v)->setFrequencyFilterType((ct::FilterType)f); }
void ctv_set_use_cuda(void* v, int on) { static_cast<ct::CtVolume*>(v)->setUseCuda(on != 0); }
void ctv_set_use_cpu_preprocessing(void* v, int on) { static_cast<ct::CtVolume*>(v)->setUseCpuPreprocessing(on != 0); }
void ctv_set_skip_projections(void* v, unsigned n) { static_cast<ct::CtVolume*>(v)->setSkipProjections(n); }
void ctv_set_spare_memory(void* v, unsigned long long mb) { static_cast<ct::CtVolume*>(v)->setGpuSpareMemory(mb); }
void ctv_reconstruct(void* v) { static_cast<ct::CtVolume*>(v)->reconstructVolume(); }
void ctv_save(void* v, const char* path, int index_order, int byte_order) {
    static_cast<ct::CtVolume*>(v)->saveVolumeToBinaryFile(path, (ct::IndexOrder)index_order, (ct::ByteOrder)byte_order);
}
void ctv_stop(void* v) { static_cast<ct::CtVolume*>(v)->stop(); }
// xMax, yMax, zMax, imageWidth, imageHeight, radius, sinogram size
void ctv_sizes(void* v, unsigned long long* out) {
    auto* c = static_cast<ct::CtVolume*>(v);
    out[0] = c->getXSize(); out[1] = c->getYSize(); out[2] = c->getZSize();
    out[3] = c->getImageWidth(); out[4] = c->getImageHeight(); out[5] = c->getReconstructionCylinderRadius();
    out[6] = c->getSinogramSize();
}
// FCD (px), pixel size, uOffset (px), base intensity
void ctv_params(void* v, float* out) {
    auto* c = static_cast<ct::CtVolume*>(v);
    out[0] = c->getFCD(); out[1] = c->getPixelSize(); out[2] = c->getUOffset(); out[3] = c->getBaseIntensity();
}
// 0 success, 1 interrupted, 2 error; message copied into msg
int ctv_status(void* v, char* msg, int len) {
    const ct::CompletionStatus s = static_cast<ct::CtVolume*>(v)->lastStatus();
    if (msg && len > 0) {
        std::strncpy(msg, s.errorMessage.c_str(), (size_t)len - 1);
        msg[len - 1] = '\0';
    }
    return s.successful ? 0 : (s.userInterrupted ? 1 : 2);
}
// getVolumeCrossSection: returns 1 and writes rows x cols floats to out (when cap, in
// floats, is large enough); 0 when there is no volume; -1 for an index out of range.
int ctv_cross_section(void* v, int axis, unsigned index, float* out, unsigned long long cap, int* rows, int* cols) {
    ct::Image im;
    try {
        im = static_cast<ct::CtVolume*>(v)->getVolumeCrossSection((ct::Axis)axis, index);
    } catch (const std::out_of_range&) {
        return -1;
    }
    *rows = im.height;
    *cols = im.width;
    if (im.empty()) return 0;
    if (out && cap >= (unsigned long long)im.width * im.height) std::memcpy(out, im.data.data(), im.data.size());
    return 1;
}
// Cross-section progress (reconstructionProgress(double, cv::Mat), Ct/CtVolume.h:273):
// fn(percentage, rows x cols floats of the cross-section at the set axis and index,
// rows, cols, user) after every finished slab; a null fn removes the callback.
typedef void (*ctv_cross_section_fn)(double, const float*, int, int, void*);
void ctv_on_cross_section_progress(void* v, ctv_cross_section_fn fn, void* user) {
    auto* c = static_cast<ct::CtVolume*>(v);
    if (!fn) {
        c->onReconstructionProgress(ct::CtVolume::CrossSectionProgressFn());
        return;
    }
    c->onReconstructionProgress(ct::CtVolume::CrossSectionProgressFn([fn, user](double pct, const ct::Image& im) {
        fn(pct, im.empty() ? nullptr : reinterpret_cast<const float*>(im.data.data()), im.height, im.width, user);
    }));
}
void ctv_set_cross_section(void* v, int axis, unsigned index) {
    auto* c = static_cast<ct::CtVolume*>(v);
    c->setCrossSectionAxis((ct::Axis)axis);
    c->setCrossSectionIndex(index);
}
// writeRawVolume on a caller's Z_FASTEST volume: 0 ok, 1 cannot open, 2 write error.
int ctv_write_raw(const float* vol, unsigned long long nx, unsigned long long ny, unsigned long long nz,
                  const char* path, int index_order, int byte_order) {
    return ct::writeRawVolume(path, vol, nx, ny, nz, (ct::IndexOrder)index_order, (ct::ByteOrder)byte_order, nullptr,
                              nullptr);
}
const float* ctv_volume(void* v, unsigned long long* n) {
    const auto& vol = static_cast<ct::CtVolume*>(v)->volumeData();
    *n = vol.size();
    return vol.data();
}
}

extern "C" {
// Decodes an image file; returns 1 on success.  depth: 0 = 8U, 1 = 16U, 2 = 32F.  The
// pixels (native byte order) are copied to out when cap is large enough.
int ctv_decode_image(const char* path, int* width, int* height, int* depth, void* out, unsigned long long cap,
                     char* err, int errlen) {
    ct::Image im;
    std::string e;
    const bool ok = ct::readImage(path, im, e);
    if (err && errlen > 0) {
        std::strncpy(err, e.c_str(), (size_t)errlen - 1);
        err[errlen - 1] = '\0';
    }
    if (!ok) return 0;
    *width = im.width;
    *height = im.height;
    *depth = (int)im.depth;
    if (out && cap >= im.data.size()) std::memcpy(out, im.data.data(), im.data.size());
    return 1;
}
}
