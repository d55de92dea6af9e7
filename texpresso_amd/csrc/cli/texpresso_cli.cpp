// texpresso -- command-line front end over the C ABI.
//
// Same commands, options, defaults and failure behaviour as texpresso's CLI
// (cli/src/main.rs:32-196):
//   texpresso compress [-o OUT] -f {bc1..bc5} [-p {speed,balanced,quality}]
//                      [--weigh-colour-by-alpha] [-w W]... INFILE
//   texpresso decompress [-o OUT] INFILE
// compress reads a PNG, encodes it on the GPU (tx_compress) and writes a DDS
// with level 0 only; decompress reads a DDS, decodes level 0 on the GPU
// (tx_decompress) and writes an RGBA8 PNG.  Profiles map to algorithms as in
// cli/src/main.rs:198-206 (speed = RangeFit, balanced = ClusterFit, quality
// = IterativeClusterFit); weights default to the perceptual ones and must
// be given exactly three times when given (104-110).
//
// Exit status: 0 on success, 2 for a malformed command line (printed with a
// usage line), 101 for a failure while running -- the statuses of the
// reference's argument parser and of a Rust panic.  JPEG input, which the
// reference decodes with the jpeg-decoder crate, is refused with status 101.
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <new>
#include <string>
#include <vector>

#include "texpresso_hip.h"

namespace {

const char* const kVersion = "texpresso_cli 2.0.2";

const char* const kTopUsage =
    "Usage: texpresso <COMMAND>\n"
    "\n"
    "Commands:\n"
    "  compress    Compress a PNG or JPEG file to DDS\n"
    "  decompress  Deompress a DDS file to PNG\n"
    "  help        Print this message or the help of the given subcommand(s)\n"
    "\n"
    "Options:\n"
    "  -h, --help     Print help\n"
    "  -V, --version  Print version\n";

const char* const kCompressUsage =
    "Compress a PNG or JPEG file to DDS\n"
    "\n"
    "Usage: texpresso compress [OPTIONS] --format <FORMAT> <INFILE>\n"
    "\n"
    "Arguments:\n"
    "  <INFILE>  Input file (PNG, JPG)\n"
    "\n"
    "Options:\n"
    "  -o, --output <OUTFILE>   Output file (DDS)\n"
    "  -f, --format <FORMAT>    Compression format [possible values: bc1, bc2, bc3, bc4, bc5]\n"
    "  -p, --profile <PROFILE>  Compressor profile (speed, balanced, quality) [default: balanced]\n"
    "                           [possible values: speed, balanced, quality]\n"
    "      --weigh-colour-by-alpha\n"
    "                           Weigh colours by alpha while fitting. Can improve perceived quality in\n"
    "                           alpha-blended images\n"
    "  -w, --weights <WEIGHTS>  Colour weights to be used for matching colours during fitting\n"
    "  -h, --help               Print help\n";

const char* const kDecompressUsage =
    "Deompress a DDS file to PNG\n"
    "\n"
    "Usage: texpresso decompress [OPTIONS] <INFILE>\n"
    "\n"
    "Arguments:\n"
    "  <INFILE>  Input file (DDS)\n"
    "\n"
    "Options:\n"
    "  -o, --output <OUTFILE>  Output file (PNG)\n"
    "  -h, --help              Print help\n";

[[noreturn]] void usage_error(const std::string& msg, const char* usage_line)
{
    fprintf(stderr, "error: %s\n\n%s\n\nFor more information, try '--help'.\n", msg.c_str(), usage_line);
    exit(2);
}

[[noreturn]] void fail(const std::string& msg)
{
    fprintf(stderr, "texpresso: %s\n", msg.c_str());
    exit(101);
}

[[noreturn]] void fail_rc(const std::string& what, int rc)
{
    fail(what + ": " + tx_error_string(rc));
}

std::vector<uint8_t> read_file(const std::string& path)
{
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) fail("Failed to open file: " + path + ": " + strerror(errno));
    std::vector<uint8_t> data;
    uint8_t buf[1 << 16];
    size_t got;
    while ((got = fread(buf, 1, sizeof(buf), f)) > 0) data.insert(data.end(), buf, buf + got);
    const bool bad = ferror(f) != 0;
    fclose(f);
    if (bad) fail("Failed to read file: " + path);
    return data;
}

void write_file(const std::string& path, const uint8_t* a, size_t na, const uint8_t* b, size_t nb)
{
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) fail("Failed to create output file: " + path + ": " + strerror(errno));
    bool ok = fwrite(a, 1, na, f) == na;
    if (ok && nb) ok = fwrite(b, 1, nb, f) == nb;
    ok = (fclose(f) == 0) && ok;
    if (!ok) fail("Failed to write output file: " + path);
}

// Path::file_name: the last component, or None for a path ending in "..".
bool file_name(std::string p, std::string& name)
{
    while (p.size() > 1 && p.back() == '/') p.pop_back();
    while (p.size() >= 2 && p.compare(p.size() - 2, 2, "/.") == 0) {
        p.resize(p.size() - 2);
        while (p.size() > 1 && p.back() == '/') p.pop_back();
    }
    const size_t slash = p.rfind('/');
    name = slash == std::string::npos ? p : p.substr(slash + 1);
    return !(name.empty() || name == "." || name == "..");
}

// Path::extension of a file name: the text after the last '.', unless that
// dot starts the name.
bool extension(const std::string& name, std::string& ext)
{
    const size_t dot = name.rfind('.');
    if (dot == std::string::npos || dot == 0) return false;
    ext = name.substr(dot + 1);
    return true;
}

// PathBuf::new().with_file_name(infile.file_name() or "output")
//           .with_extension(ext)   (cli/src/main.rs:123-127, 168-172)
std::string default_output(const std::string& infile, const char* ext)
{
    std::string name;
    if (!file_name(infile, name)) name = "output";
    const size_t dot = name.rfind('.');
    if (dot != std::string::npos && dot != 0) name.resize(dot);
    return name + "." + ext;
}

// 4 * w * h bytes of RGBA, refusing sizes whose product overflows or that no
// texture the library accepts can have (sides below 2^31).
size_t rgba_bytes(size_t w, size_t h)
{
    if (w == 0 || h == 0 || w > 0x7FFFFFFFu || h > 0x7FFFFFFFu || w > (SIZE_MAX / 4) / h)
        fail("Image size out of range: " + std::to_string(w) + "x" + std::to_string(h));
    return 4 * w * h;
}

std::string lower(std::string s)
{
    for (char& c : s)
        if (c >= 'A' && c <= 'Z') c = (char)(c - 'A' + 'a');
    return s;
}

// f32::from_str: decimal or inf/infinity/nan, no surrounding space, no hex.
bool parse_f32(const char* s, float* out)
{
    if (!*s) return false;
    for (const char* c = s; *c; ++c)
        if (*c == 'x' || *c == 'X' || *c == ' ' || *c == '\t' || *c == '\n' || *c == '(') return false;
    char* end = nullptr;
    errno = 0;
    const float v = strtof(s, &end);
    if (*end != '\0') return false;
    *out = v;  // overflow gives +-inf and underflow a (sub)normal or zero, as in Rust
    return true;
}

struct Args {
    std::vector<std::string> pos;
    std::string output, format, profile;
    bool have_output = false, have_format = false, have_profile = false, weigh_alpha = false;
    std::vector<float> weights;
};

// The option forms clap accepts: "-o X", "-oX", "--output X", "--output=X";
// "--" ends the options.
Args parse_args(int argc, char** argv, bool compress, const char* usage_line, const char* help)
{
    Args a;
    bool opts = true;
    for (int i = 0; i < argc; ++i) {
        std::string s = argv[i];
        if (opts && s == "--") {
            opts = false;
            continue;
        }
        if (!opts || s.size() < 2 || s[0] != '-') {
            a.pos.push_back(s);
            continue;
        }
        std::string key, val;
        bool inline_val = false;
        if (s[1] == '-') {
            const size_t eq = s.find('=');
            key = s.substr(2, eq == std::string::npos ? std::string::npos : eq - 2);
            if (eq != std::string::npos) {
                val = s.substr(eq + 1);
                inline_val = true;
            }
        } else {
            const char k = s[1];
            static const char* const shorts[][2] = {{"o", "output"}, {"f", "format"}, {"p", "profile"},
                                                    {"w", "weights"}, {"h", "help"}};
            key.clear();
            for (auto& m : shorts)
                if (k == m[0][0]) key = m[1];
            if (key.empty()) usage_error(std::string("unexpected argument '") + s + "' found", usage_line);
            if (s.size() > 2) {
                val = s.substr(s[2] == '=' ? 3 : 2);
                inline_val = true;
            }
        }
        if (key == "help") {
            fputs(help, stdout);
            exit(0);
        }
        const bool takes_value =
            key == "output" || (compress && (key == "format" || key == "profile" || key == "weights"));
        if (key == "weigh-colour-by-alpha" && compress) {
            if (inline_val) usage_error("unexpected value '" + val + "' for '--weigh-colour-by-alpha' found", usage_line);
            if (a.weigh_alpha)
                usage_error("the argument '--weigh-colour-by-alpha' cannot be used multiple times", usage_line);
            a.weigh_alpha = true;
            continue;
        }
        if (!takes_value) usage_error("unexpected argument '" + s + "' found", usage_line);
        if (!inline_val) {
            if (i + 1 >= argc) usage_error("a value is required for '--" + key + "' but none was supplied", usage_line);
            val = argv[++i];
        }
        if (key == "output") {
            if (a.have_output) usage_error("the argument '--output <OUTFILE>' cannot be used multiple times", usage_line);
            a.output = val;
            a.have_output = true;
        } else if (key == "format") {
            if (a.have_format) usage_error("the argument '--format <FORMAT>' cannot be used multiple times", usage_line);
            a.format = val;
            a.have_format = true;
        } else if (key == "profile") {
            if (a.have_profile)
                usage_error("the argument '--profile <PROFILE>' cannot be used multiple times", usage_line);
            a.profile = val;
            a.have_profile = true;
        } else {
            float w;
            if (!parse_f32(val.c_str(), &w))
                usage_error("invalid value '" + val + "' for '--weights <WEIGHTS>': invalid float literal", usage_line);
            a.weights.push_back(w);
        }
    }
    if (a.pos.empty()) usage_error("the following required arguments were not provided:\n  <INFILE>", usage_line);
    if (a.pos.size() > 1) usage_error("unexpected argument '" + a.pos[1] + "' found", usage_line);
    return a;
}

int run_compress(int argc, char** argv)
{
    const char* usage_line = "Usage: texpresso compress [OPTIONS] --format <FORMAT> <INFILE>";
    Args a = parse_args(argc, argv, true, usage_line, kCompressUsage);
    if (!a.have_format)
        usage_error("the following required arguments were not provided:\n  --format <FORMAT>", usage_line);
    static const char* const formats[5] = {"bc1", "bc2", "bc3", "bc4", "bc5"};
    int format = -1;
    for (int f = 0; f < 5; ++f)
        if (a.format == formats[f]) format = f;
    if (format < 0)
        usage_error("invalid value '" + a.format + "' for '--format <FORMAT>'\n  [possible values: bc1, bc2, bc3, bc4, bc5]",
                    usage_line);
    int algorithm = TX_CLUSTER_FIT;  // profile "balanced"
    if (a.have_profile) {
        if (a.profile == "speed") algorithm = TX_RANGE_FIT;
        else if (a.profile == "balanced") algorithm = TX_CLUSTER_FIT;
        else if (a.profile == "quality") algorithm = TX_ITERATIVE_CLUSTER_FIT;
        else
            usage_error("invalid value '" + a.profile +
                            "' for '--profile <PROFILE>'\n  [possible values: speed, balanced, quality]",
                        usage_line);
    }
    tx_params params = tx_params_default();
    params.algorithm = algorithm;
    if (!a.weights.empty()) {
        if (a.weights.size() != 3) fail("Weights must have 3 values");
        for (int c = 0; c < 3; ++c) params.weights[c] = a.weights[c];
    }
    params.weigh_colour_by_alpha = a.weigh_alpha ? 1 : 0;

    const std::string& infile = a.pos[0];
    const std::string outfile = a.have_output ? a.output : default_output(infile, "dds");
    std::string name, ext;
    if (!file_name(infile, name) || !extension(name, ext)) fail("Input filename has no extension, can't guess type");
    ext = lower(ext);
    if (ext == "jpg" || ext == "jpeg") fail("JPEG input is not supported by this build; convert it to PNG");
    if (ext != "png") fail("Unrecognized image format. Supported formats are PNG and JPEG");

    const std::vector<uint8_t> file = read_file(infile);
    size_t w = 0, h = 0;
    int rc = tx_png_info(file.data(), file.size(), &w, &h);
    if (rc != TX_OK) fail_rc("Failed to read PNG header. Is this really a PNG file?", rc);
    std::vector<uint8_t> rgba(rgba_bytes(w, h));
    rc = tx_png_decode(file.data(), file.size(), rgba.data(), rgba.size());
    if (rc != TX_OK) fail_rc("Failed to decode PNG image data", rc);

    std::vector<uint8_t> blocks(tx_compressed_size(format, w, h));
    rc = tx_compress(format, rgba.data(), rgba.size(), w, h, &params, blocks.data(), blocks.size());
    if (rc != TX_OK) fail_rc("compress failed", rc);
    std::vector<uint8_t> header(tx_dds_header_size());
    rc = tx_dds_write_header(format, w, h, 1, header.data(), header.size());
    if (rc < 0) fail_rc("Failed to build the DDS header", rc);
    write_file(outfile, header.data(), header.size(), blocks.data(), blocks.size());
    return 0;
}

int run_decompress(int argc, char** argv)
{
    const char* usage_line = "Usage: texpresso decompress [OPTIONS] <INFILE>";
    Args a = parse_args(argc, argv, false, usage_line, kDecompressUsage);
    const std::string& infile = a.pos[0];
    const std::string outfile = a.have_output ? a.output : default_output(infile, "png");
    const std::vector<uint8_t> file = read_file(infile);
    int format = 0;
    size_t w = 0, h = 0, off = 0;
    uint32_t levels = 0;
    int rc = tx_dds_parse_header(file.data(), file.size(), &format, &w, &h, &levels, &off);
    if (rc != TX_OK) fail_rc("Failed to read the DDS file", rc);
    std::vector<uint8_t> rgba(rgba_bytes(w, h));
    rc = tx_decompress(format, file.data() + off, file.size() - off, w, h, rgba.data(), rgba.size());
    if (rc != TX_OK) fail_rc("decompress failed", rc);
    std::vector<uint8_t> png(tx_png_encode_bound(w, h));
    size_t len = 0;
    if (png.empty()) fail("Image too large for PNG output");
    rc = tx_png_encode(rgba.data(), w, h, png.data(), png.size(), &len);
    if (rc != TX_OK) fail_rc("Failed to encode the PNG file", rc);
    write_file(outfile, png.data(), len, nullptr, 0);
    return 0;
}

}  // namespace

int run(int argc, char** argv)
{
    if (argc < 2) {
        fputs(kTopUsage, stderr);
        return 2;
    }
    const std::string cmd = argv[1];
    if (cmd == "compress") return run_compress(argc - 2, argv + 2);
    if (cmd == "decompress") return run_decompress(argc - 2, argv + 2);
    if (cmd == "-V" || cmd == "--version") {
        printf("%s\n", kVersion);
        return 0;
    }
    if (cmd == "-h" || cmd == "--help") {
        fputs(kTopUsage, stdout);
        return 0;
    }
    if (cmd == "help") {
        const std::string sub = argc > 2 ? argv[2] : "";
        fputs(sub == "compress" ? kCompressUsage : sub == "decompress" ? kDecompressUsage : kTopUsage, stdout);
        return 0;
    }
    usage_error("unrecognized subcommand '" + cmd + "'", "Usage: texpresso <COMMAND>");
}

int main(int argc, char** argv)
{
    try {
        return run(argc, argv);
    } catch (const std::bad_alloc&) {
        fail("Out of host memory");
    }
}
