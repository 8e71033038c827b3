// Command-line surface of the driver.  The observable behaviour is the
// reference's (src/Config.hpp:7-92): the same flags and defaults, the same
// usage text and error messages, exit status 1 after an error and 0 after
// -h, the same echo lines and the same default output name
// <image dir>/<image stem>_<version>.hdr.  Two flags are added: --debug and
// -c/--channels.
//
// Design: every option is one row of a table (spellings, value placeholder,
// help text, and a setter that validates and stores the value).  The usage
// text is printed from the table, and one loop looks each argument up in it.
#ifndef LPA_ICI_CONFIG_HPP
#define LPA_ICI_CONFIG_HPP

#include <cstdlib>
#include <iostream>
#include <sstream>
#include <string>

namespace lpa_cli {

// What the driver runs with.
struct Settings {
    std::string image = "data/barbara.png";
    std::string kernels = "kernels/lpa_kernels_m_1_0_d_16_h12_1_2_4_8_16_24_32_symmetric.txt";
    std::string output;              // filled from the image path when not given
    float sigma = 20.0f / 255.0f;    // on the [0,1] scale (the flag takes the 0-255 scale)
    unsigned int seed = 20250910u;
    float gamma = 2.0f;
    std::string debug_dir;           // non-empty: per-direction debugging outputs go there
    int channels = 1;                // 1: gray (the reference's only mode); 3: RGB
};

// Path helpers.  folder("a/b/c.png") = "a/b/" (empty without a slash);
// stem("a/b/c.d.png") = "c.d" (the file name up to its last dot).
inline std::string folder(const std::string &path)
{
    const std::string::size_type cut = path.rfind('/');
    return cut == std::string::npos ? std::string() : path.substr(0, cut + 1);
}
inline std::string stem(const std::string &path)
{
    const std::string file = path.substr(folder(path).size());
    return file.substr(0, file.rfind('.'));
}
// A file named `name` in the folder that holds `path`.
inline std::string beside(const std::string &path, const std::string &name) { return folder(path) + name; }

// A rejected value: the message printed after "Error: ".
struct UsageError {
    std::string message;
};

// Value parsers shared by the setters.
inline float positive_number(const std::string &flag, const std::string &text)
{
    char *end = nullptr;
    const float v = std::strtof(text.c_str(), &end);
    if (text.empty() || *end != '\0' || !(v > 0.0f))
        throw UsageError{flag + " needs a positive number, got '" + text + "'"};
    return v;
}
inline unsigned int unsigned_number(const std::string &flag, const std::string &text)
{
    char *end = nullptr;
    const unsigned long v = std::strtoul(text.c_str(), &end, 10);
    if (text.empty() || *end != '\0') throw UsageError{flag + " needs an unsigned integer, got '" + text + "'"};
    return static_cast<unsigned int>(v);
}

// One row per option.  `shortname` is "" when there is none; `setter` is
// null for -h.  `help` receives the defaults.
struct Option {
    const char *shortname;
    const char *longname;
    const char *placeholder;
    std::string (*help)(const Settings &defaults);
    void (*setter)(Settings &s, const std::string &flag, const std::string &value);
};

template <typename T>
inline std::string text_of(const T &v)
{
    std::ostringstream os;
    os << v;
    return os.str();
}

inline const Option *option_table(int *count)
{
    static const Option table[] = {
        {"-i", "--image", "<file>", [](const Settings &d) { return "input image (default " + d.image + ")"; },
         [](Settings &s, const std::string &, const std::string &v) { s.image = v; }},
        {"-k", "--kernels", "<file>", [](const Settings &d) { return "LPA kernel file (default " + d.kernels + ")"; },
         [](Settings &s, const std::string &, const std::string &v) { s.kernels = v; }},
        {"-o", "--output", "<file>",
         [](const Settings &) { return std::string("denoised .hdr output (default <image dir>/<image name>_<version>.hdr)"); },
         [](Settings &s, const std::string &, const std::string &v) { s.output = v; }},
        {"-s", "--sigma", "<value>",
         [](const Settings &) { return std::string("noise standard deviation on the 0-255 scale (default 20)"); },
         [](Settings &s, const std::string &f, const std::string &v) { s.sigma = positive_number(f, v) / 255.0f; }},
        {"-g", "--gamma", "<value>", [](const Settings &d) { return "ICI threshold (default " + text_of(d.gamma) + ")"; },
         [](Settings &s, const std::string &f, const std::string &v) { s.gamma = positive_number(f, v); }},
        {"", "--seed", "<value>", [](const Settings &d) { return "noise seed (default " + text_of(d.seed) + ")"; },
         [](Settings &s, const std::string &f, const std::string &v) { s.seed = unsigned_number(f, v); }},
        {"", "--debug", "<dir>",
         [](const Settings &) {
             return std::string("also write each direction's estimate and scale map as .hdr in <dir>");
         },
         [](Settings &s, const std::string &, const std::string &v) { s.debug_dir = v; }},
        {"-c", "--channels", "<1|3>",
         [](const Settings &) { return std::string("1: gray (default); 3: RGB, every channel denoised on its own"); },
         [](Settings &s, const std::string &f, const std::string &v) {
             if (v != "1" && v != "3") throw UsageError{f + " needs 1 or 3, got '" + v + "'"};
             s.channels = v == "3" ? 3 : 1;
         }},
        {"-h", "--help", "", [](const Settings &) { return std::string("show this message"); }, nullptr},
    };
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}

// "  -i, --image <file>     input image (...)": spellings and placeholder in
// a 23-column field after two spaces, then the help text.
inline void print_usage(const char *prog)
{
    const Settings defaults;
    int n = 0;
    const Option *table = option_table(&n);
    std::cout << "Usage: " << prog << " [options]\n";
    for (int i = 0; i < n; ++i) {
        const Option &o = table[i];
        std::string left = (*o.shortname ? std::string(o.shortname) + ", " : std::string(4, ' ')) + o.longname;
        if (*o.placeholder) left += std::string(" ") + o.placeholder;
        if (left.size() < 23) left.resize(23, ' ');
        std::cout << "  " << left << o.help(defaults) << "\n";
    }
}

inline const Option *find_option(const std::string &arg)
{
    int n = 0;
    const Option *table = option_table(&n);
    for (int i = 0; i < n; ++i)
        if (arg == table[i].longname || (*table[i].shortname && arg == table[i].shortname)) return &table[i];
    return nullptr;
}

// Reads argv into Settings.  Every option but -h takes exactly one value.
// On a bad argument: "Error: <message>", a blank line and the usage text,
// then exit status 1; after -h the usage text and exit status 0.  Fills the
// default output name and echoes the settings.
inline Settings parse_command_line(int argc, char **argv, const std::string &version)
{
    Settings s;
    auto reject = [&](const std::string &message) {
        std::cerr << "Error: " << message << "\n\n";
        print_usage(argv[0]);
        std::exit(EXIT_FAILURE);
    };
    int i = 1;
    while (i < argc) {
        const std::string flag = argv[i];
        const Option *o = find_option(flag);
        if (o && !o->setter) {
            print_usage(argv[0]);
            std::exit(EXIT_SUCCESS);
        }
        if (i + 1 == argc) reject("unknown option or missing value for '" + flag + "'");
        if (!o) reject("unknown option '" + flag + "'");
        try {
            o->setter(s, flag, argv[i + 1]);
        } catch (const UsageError &e) {
            reject(e.message);
        }
        i += 2;
    }
    if (s.output.empty()) s.output = folder(s.image) + stem(s.image) + "_" + version + ".hdr";
    std::cout << "Image: " << s.image << "\n"
              << "Kernels: " << s.kernels << "\n"
              << "Output: " << s.output << "\n"
              << "Sigma: " << s.sigma * 255.0f << "/255, gamma: " << s.gamma << ", seed: " << s.seed << "\n";
    return s;
}

}  // namespace lpa_cli

// The reference's names for the same settings and parser (src/Config.hpp:7-14,
// 22-24, 39), so a main() written against them builds unchanged.
struct Config {
    std::string image_file, kernel_file, output_file;
    float sigma_noise;
    unsigned int noise_seed;
    float ici_gamma;
};

inline Config parse_args(int argc, char **argv, const std::string &version)
{
    const lpa_cli::Settings s = lpa_cli::parse_command_line(argc, argv, version);
    return Config{s.image, s.kernels, s.output, s.sigma, s.seed, s.gamma};
}

inline std::string sibling_path(const std::string &path, const std::string &name) { return lpa_cli::beside(path, name); }

#endif
