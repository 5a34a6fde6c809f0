// Command-line front end with Ct's console flags (Ct/main.cpp:53-234): load the CSV
// geometry file, reconstruct on the GPU through the CtVolume facade, write the volume.
// Differences from the reference: there is no GUI mode (running without arguments
// prints the help), -n (CPU reconstruction) is reported as unavailable, and the exit
// status is non-zero when loading, reconstruction or saving fails.
#include <algorithm>
#include <cctype>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "ctvolume.h"

namespace {

void print_help() {
    std::cout << "Usage: ct [parameters]\n"
                 "Parameters:\n"
                 "\t-i [path]\tFile path to the input config file. Long: --input.\n"
                 "\t-o [path]\tFile path for the output file. Long: --output.\n"
                 "\t-b \t\tOptional. Run with background priority. Long: --background.\n"
                 "\t-f [option] \tOptional. Sets the preprocessing filter. Options are\n"
                 "\t\t\t'ramlak', 'shepplogan' and 'hann'. Long: --filter.\n\t\t\tDefault: ramlak\n"
                 "\t-n \t\tOptional. Disables GPU reconstruction (not available in ct_amd). Long: --nocuda.\n"
                 "\t-c \t\tOptional. Uses the CPU preprocessing formulas. Long: --cpuPreprocessing.\n"
                 "\t-d 0,1,..,n \tOptional. Sets the GPU devices that shall be used (the first is used).\n"
                 "\t\t\tLong: --cudadevices. Default: 0\n"
                 "\t-w x,y \tSpecifies the coefficients for the GPU weights. Long: --weights. Default: 1,1\n"
                 "\t-m [number] \tOptional. Sets the amount of VRAM to spare in Mb. Long: --cudasparememory.\n"
                 "\t\t\tDefault: 200\n"
                 "\t-e [option] \tOptional. Sets the byte order of the output. Options\n"
                 "\t\t\tare 'littleendian' and 'bigendian'. Long: --byteorder. Default: littleendian\n"
                 "\t-j [option] \tOptional. Sets the index order of the output.\n"
                 "\t\t\tOptions are 'zfastest' and 'xfastest'. Long: --indexorder.\n"
                 "\t-h \t\tDisplay this help. Long: --help.\n"
                 "\t--xmin 0..1\tOptional. The lower x bound of the part of the volume\n"
                 "\t\t\tthat will be reconstructed as float between 0 and 1.\n"
                 "\t--xmax, --ymin, --ymax, --zmin, --zmax 0..1\tLikewise.\n";
}

std::string lower(std::string s) {
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return s;
}

std::vector<std::string> split_commas(const std::string& s) {
    std::vector<std::string> out;
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, ',')) out.push_back(item);
    return out;
}

}  // namespace

int main(int argc, char* argv[]) {
    if (argc < 2 || std::string(argv[1]) == "--help" || std::string(argv[1]) == "-h") {
        print_help();
        return 0;
    }
    ct::CtVolume volume;
    bool inputProvided = false, outputProvided = false;
    bool useCuda = ct::CtVolume::cudaAvailable();
    bool useCpuPreprocessing = false;
    ct::FilterType filterType = ct::FilterType::RAMLAK;
    // Ct/main.cpp:62: the console default is X fastest; "-j zfastest" selects Z fastest
    ct::IndexOrder indexOrder = ct::IndexOrder::X_FASTEST;
    ct::ByteOrder byteOrder = ct::ByteOrder::LittleEndian;
    std::string filterTypeString = "Ram-Lak";
    std::string input, output;
    double xmin = 0, xmax = 1, ymin = 0, ymax = 1, zmin = 0, zmax = 1;
    auto parse_double = [&](int& i, double& out) {
        if (++i < argc) out = std::stod(argv[i]);
    };
    try {
        for (int i = 1; i < argc; ++i) {
            const std::string a = argv[i];
            if (a == "--xmin") parse_double(i, xmin);
            else if (a == "--xmax") parse_double(i, xmax);
            else if (a == "--ymin") parse_double(i, ymin);
            else if (a == "--ymax") parse_double(i, ymax);
            else if (a == "--zmin") parse_double(i, zmin);
            else if (a == "--zmax") parse_double(i, zmax);
            else if (a == "-f" || a == "--filter") {
                if (++i < argc) {
                    const std::string f = lower(argv[i]);
                    if (f == "shepplogan") {
                        filterType = ct::FilterType::SHEPP_LOGAN;
                        filterTypeString = "Shepp-Logan";
                    } else if (f == "hann") {
                        filterType = ct::FilterType::HANN;
                        filterTypeString = "Hann";
                    }
                }
            } else if (a == "-i" || a == "--input") {
                if (++i < argc) {
                    input = argv[i];
                    inputProvided = true;
                }
            } else if (a == "-o" || a == "--output") {
                if (++i < argc) {
                    output = argv[i];
                    outputProvided = true;
                }
            } else if (a == "-b" || a == "--background") {
                // background priority: accepted, no effect on Linux (as in the reference)
            } else if (a == "-n" || a == "--nocuda") {
                useCuda = false;
            } else if (a == "-c" || a == "--cpuPreprocessing") {
                useCpuPreprocessing = true;
            } else if (a == "-d" || a == "--cudadevices") {
                if (++i < argc) {
                    std::vector<int> devices;
                    for (auto& d : split_commas(argv[i])) devices.push_back(std::atoi(d.c_str()));
                    if (devices.empty()) {
                        std::cout << "When using the flag -d you have to specify at least one device ID." << std::endl;
                        return 1;
                    }
                    volume.setActiveCudaDevices(devices);
                }
            } else if (a == "-m" || a == "--cudasparememory") {
                double spare = 200;
                parse_double(i, spare);
                volume.setGpuSpareMemory((size_t)spare);
            } else if (a == "-e" || a == "--byteorder") {
                if (++i < argc && lower(argv[i]) == "bigendian") byteOrder = ct::ByteOrder::BigEndian;
            } else if (a == "-j" || a == "--indexorder") {
                if (++i < argc && lower(argv[i]) == "zfastest") indexOrder = ct::IndexOrder::Z_FASTEST;
            } else if (a == "-w" || a == "--weights") {
                if (++i < argc) {
                    const auto c = split_commas(argv[i]);
                    if (c.size() != 2) {
                        std::cout << "When using the flag -w you have to specify exactly two coefficients." << std::endl;
                        return 1;
                    }
                    volume.setGpuCoefficients(std::atof(c[0].c_str()), std::atof(c[1].c_str()));
                }
            } else {
                std::cout << "Unknown or misplaced parameter " << a << "." << std::endl;
                return 1;
            }
        }
    } catch (const std::exception&) {
        std::cout << "Could not parse a numeric parameter." << std::endl;
        return 1;
    }
    if (!(inputProvided && outputProvided)) {
        std::cout << "You must provide a file path to the config file as input as well as a file path for the output."
                  << std::endl;
        return 1;
    }
    const std::vector<std::string> names = volume.getCudaDeviceList();
    const std::vector<int> active = volume.getActiveCudaDevices();
    std::cout << std::endl << "Beginning reconstruction." << std::endl;
    std::cout << "\tInput:\t\t\t" << input << std::endl;
    std::cout << "\tOutput:\t\t\t" << output << std::endl;
    std::cout << "\tFilter type:\t\t" << filterTypeString << std::endl;
    std::cout << "\tUsing GPU:\t\t" << (useCuda ? "YES" : "NO") << std::endl;
    std::cout << "\tUsing CPU preprocessing formulas: " << (useCpuPreprocessing ? "YES" : "NO") << std::endl;
    if (!names.empty()) {
        std::cout << "\tGPU devices:" << std::endl;
        for (size_t i = 0; i < names.size(); ++i) {
            const bool on = std::find(active.begin(), active.end(), (int)i) != active.end() && useCuda;
            std::cout << "\t\t[" << (on ? "X" : " ") << "] " << names[i] << std::endl;
        }
    }
    std::cout << "\tVolume bounds:\t\tx: [" << xmin << " .. " << xmax << "]" << std::endl;
    std::cout << "\t\t\t\ty: [" << ymin << " .. " << ymax << "]" << std::endl;
    std::cout << "\t\t\t\tz: [" << zmin << " .. " << zmax << "]" << std::endl;
    std::cout << "\tByte order:\t\t" << (byteOrder == ct::ByteOrder::LittleEndian ? "Little endian" : "Big endian")
              << std::endl;
    std::cout << "\tIndex order:\t\t" << (indexOrder == ct::IndexOrder::Z_FASTEST ? "Z fastest" : "X fastest")
              << std::endl
              << std::endl;
    volume.setEmitSignals(false);
    bool ok = volume.sinogramFromImages(input);
    volume.setVolumeBounds((float)xmin, (float)xmax, (float)ymin, (float)ymax, (float)zmin, (float)zmax);
    std::cout << std::endl << "The resulting volume dimensions will be:" << std::endl << std::endl
              << "\t" << volume.getXSize() << "x" << volume.getYSize() << "x" << volume.getZSize() << " (x:y:z)"
              << std::endl << std::endl;
    volume.setFrequencyFilterType(filterType);
    volume.setUseCuda(useCuda);
    volume.setUseCpuPreprocessing(useCpuPreprocessing);
    volume.reconstructVolume();
    ok = ok && volume.lastStatus().successful;
    volume.saveVolumeToBinaryFile(output, indexOrder, byteOrder);
    ok = ok && volume.lastStatus().successful;
    return ok ? 0 : 1;
}
