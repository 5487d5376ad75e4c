// output.cpp -- output files in the reference's format
// (GeodesicOutputHandler, FOORT/InputOutput.cpp:62-331): one file per
// diagnostic named prefix_[timestamp_]Name[_n][.ext], an optional first line
// describing the run, then one line "row col  value" per geodesic in the order
// the mesh handed the geodesics out.  Geodesics are cached and written when the
// cache would overflow and at the end; a file holds at most `per_file`
// geodesics, after which numbered files follow.
#include "host.hpp"

#include <algorithm>
#include <charconv>
#include <cstdio>
#include <memory>
#include <filesystem>
#include <iostream>
#include <limits>

namespace foort_host {

namespace {
// the cache limit the reference can hold: min(requested, vector::max_size() - 1, ULONG_MAX - 1)
unsigned long cache_limit(unsigned long n)
{
    const unsigned long vmax =
        static_cast<unsigned long>(std::numeric_limits<std::ptrdiff_t>::max()) / sizeof(std::vector<std::string>);
    return std::min({n, vmax - 1, std::numeric_limits<unsigned long>::max() - 1});
}
} // namespace

OutputHandler::OutputHandler(std::string prefix, std::string timestamp, std::string extension,
                             std::vector<std::string> diag_names, unsigned long to_cache, unsigned long per_file,
                             std::string first_line)
    : m_prefix(std::move(prefix)), m_timestamp(std::move(timestamp)), m_extension(std::move(extension)),
      m_names(std::move(diag_names)), m_to_cache(cache_limit(to_cache)), m_per_file(per_file),
      m_first_line_on(!first_line.empty()), m_first_line(std::move(first_line))
{
    if (m_prefix.empty() || m_per_file == 0)
        m_console = true;
}

std::string OutputHandler::getFullDescriptionStr() const
{
    if (!m_console)
        return "Output Handler: Basic (value diagnostic) file name: " + FileName(0, 1) +
               ", caching outputs: " + std::to_string(m_to_cache) +
               ", geodesics per file: " + std::to_string(m_per_file) +
               ", printing first line info: " + std::to_string(m_first_line_on ? 1 : 0);
    return "Output Handler: Writing to console, caching outputs: " + std::to_string(m_to_cache);
}

void OutputHandler::PrepareForOutput(uint64_t count)
{
    if (m_cache.size() + count > m_to_cache)
        WriteCached();
    m_prev_cached = m_cache.size();
    m_cache.resize(m_cache.size() + count, Entry{0, 0, {0, 0, 0, 0}, kNoText});
}

void OutputHandler::NewGeodesicOutput(uint64_t index, uint64_t row, uint64_t col, std::vector<std::string> texts)
{
    Entry &e = m_cache[m_prev_cached + index];
    e.row = row;
    e.col = col;
    e.text = m_texts.size();
    m_texts.push_back(std::move(texts));
}

void OutputHandler::NewGeodesicValues(uint64_t index, uint64_t row, uint64_t col, const double *values)
{
    Entry &e = m_cache[m_prev_cached + index];
    e.row = row;
    e.col = col;
    for (size_t d = 0; d < m_kinds.size(); ++d)
        e.values[d] = values[d];
    e.text = kNoText;
}

namespace {
// "row col  value\n" for entries of values: std::to_string's text of an int or of a double ("%f")
void append_line(std::string &b, uint64_t row, uint64_t col, double v, char kind)
{
    char num[512];
    b.append(num, std::to_chars(num, num + 48, row).ptr);
    b += ' ';
    b.append(num, std::to_chars(num, num + 48, col).ptr);
    b += "  ";
    if (kind == 1)
        b.append(num, std::to_chars(num, num + sizeof num, static_cast<int>(v)).ptr);
    else
        b.append(num, std::to_chars(num, num + sizeof num, v, std::chars_format::fixed, 6).ptr);
    b += '\n';
}
} // namespace

void OutputHandler::FormatAhead()
{
    const uint64_t lo = m_ahead_end, hi = m_cache.size();
    if (m_console || hi <= lo || m_kinds.empty())
        return;
    for (uint64_t j = lo; j < hi; ++j)
        if (m_cache[j].text != kNoText)
            return; // texts: written by WriteCached itself
    struct Line {
        uint64_t row, col;
        double values[kMaxValues];
    };
    auto lines = std::make_shared<std::vector<Line>>(hi - lo);
    for (uint64_t j = lo; j < hi; ++j) {
        const Entry &e = m_cache[j];
        Line &l = (*lines)[j - lo];
        l.row = e.row;
        l.col = e.col;
        std::copy(e.values, e.values + kMaxValues, l.values);
    }
    const std::vector<char> kinds = m_kinds;
    m_ahead.push_back(Ahead{lo, hi, std::async(std::launch::async, [lines, kinds] {
                                std::vector<std::string> out(kinds.size());
                                for (size_t d = 0; d < kinds.size(); ++d) {
                                    out[d].reserve(lines->size() * 24);
                                    for (const Line &l : *lines)
                                        append_line(out[d], l.row, l.col, l.values[d], kinds[d]);
                                }
                                return out;
                            })});
    m_ahead_end = hi;
}

void OutputHandler::AppendText(std::string &b, const Entry &e, int d) const
{
    if (e.text != kNoText) {
        b += m_texts[e.text][d];
        return;
    }
    char num[512];
    if (m_kinds[d] == kInt) {
        b.append(num, std::to_chars(num, num + sizeof num, static_cast<int>(e.values[d])).ptr);
    } else {
        // what std::to_string(double) writes ("%f"): correctly rounded to 6 decimals
        auto r = std::to_chars(num, num + sizeof num, e.values[d], std::chars_format::fixed, 6);
        b.append(num, r.ptr);
    }
}

void OutputHandler::OutputFinished() { WriteCached(); }

std::string OutputHandler::FileName(int diag, unsigned short filenr) const
{
    std::string name = m_prefix + "_";
    if (!m_timestamp.empty())
        name += m_timestamp + "_";
    name += m_names[diag];
    if (filenr > 1)
        name += "_" + std::to_string(filenr);
    if (!m_extension.empty())
        name += "." + m_extension;
    return name;
}

void OutputHandler::OpenForFirstTime(const std::string &name)
{
    auto slash = name.find_last_of('/');
    if (slash != std::string::npos) {
        std::error_code ec;
        std::filesystem::create_directories(name.substr(0, slash), ec);
    }
    std::FILE *f = std::fopen(name.c_str(), "wb");
    if (!f) {
        screen_output("Output file error! Could not open " + name + ". Will write rest of output to console.",
                      Level::Warning);
        m_console = true;
        return;
    }
    if (m_first_line_on) {
        std::fputs(m_first_line.c_str(), f);
        std::fputc('\n', f);
    }
    std::fclose(f);
}

// WriteCachedOutputToFile, FOORT/InputOutput.cpp:129-267
void OutputHandler::WriteCached()
{
    // Lines formatted ahead cover a prefix of the cache.  They are used when the rest can be
    // formatted the same way (values, no texts) and every line of this write goes to the
    // current files (no split into a next file); otherwise the lines are formatted below.
    bool values_only = true;
    for (uint64_t j = m_ahead_end; j < m_cache.size() && values_only; ++j)
        values_only = m_cache[j].text == kNoText;
    if (!m_cache.empty() && !m_console && !m_ahead.empty() && values_only && m_in_file + m_cache.size() <= m_per_file) {
        {
            screen_output("Writing cached geodesic output to file(s)...", Level::Subproc);
            const int ndiag = static_cast<int>(m_kinds.size());
            std::vector<std::vector<std::string>> texts;
            for (Ahead &a : m_ahead)
                texts.push_back(a.text.get());
            // the entries not formatted ahead, in parallel chunks
            const uint64_t lo = m_ahead_end, n = m_cache.size() - lo, kChunk = 1u << 16;
            const long long chunks = static_cast<long long>((n + kChunk - 1) / kChunk);
            std::vector<std::vector<std::string>> rest(static_cast<size_t>(chunks),
                                                       std::vector<std::string>(static_cast<size_t>(ndiag)));
#pragma omp parallel for schedule(dynamic, 1)
            for (long long c = 0; c < chunks; ++c) {
                const uint64_t a = lo + static_cast<uint64_t>(c) * kChunk, b = std::min<uint64_t>(a + kChunk, lo + n);
                for (int d = 0; d < ndiag; ++d) {
                    std::string &t = rest[static_cast<size_t>(c)][static_cast<size_t>(d)];
                    t.reserve((b - a) * 24);
                    for (uint64_t j = a; j < b; ++j)
                        append_line(t, m_cache[j].row, m_cache[j].col, m_cache[j].values[d], m_kinds[d]);
                }
            }
            for (auto &r : rest)
                texts.push_back(std::move(r));
            m_ahead.clear();
            m_ahead_end = 0;
            if (m_in_file == 0)
                for (int i = 0; i < ndiag && !m_console; ++i)
                    OpenForFirstTime(FileName(i, static_cast<unsigned short>(m_full_files + 1)));
            for (int d = 0; d < ndiag && !m_console; ++d) {
                const std::string name = FileName(d, static_cast<unsigned short>(m_full_files + 1));
                std::FILE *f = std::fopen(name.c_str(), "ab");
                if (!f) {
                    screen_output("Output file error! Could not open " + name +
                                      ". Will write rest of output to console.",
                                  Level::Warning);
                    m_console = true;
                    break;
                }
                for (const auto &t : texts)
                    std::fwrite(t[static_cast<size_t>(d)].data(), 1, t[static_cast<size_t>(d)].size(), f);
                std::fclose(f);
            }
            if (!m_console) {
                // the bookkeeping of the general path below for a write into one file
                unsigned long last_count = m_cache.size();
                if (last_count == m_per_file) {
                    ++m_full_files;
                    last_count = 0;
                }
                m_in_file += last_count;
                screen_output("Done writing cached geodesic output to file(s).", Level::Subproc);
                m_cache.clear();
                m_texts.clear();
                return;
            }
            // a file could not be opened: everything goes to the console below
        }
    }
    for (Ahead &a : m_ahead)
        a.text.wait(); // not used: finish reading the copies before they go
    m_ahead.clear();
    m_ahead_end = 0;
    if (m_cache.empty())
        return;
    if (!m_console) {
        screen_output("Writing cached geodesic output to file(s)...", Level::Subproc);
        const unsigned long total = m_cache.size();
        unsigned short nfiles = 1;
        while (m_in_file + total > static_cast<unsigned long>(nfiles) * m_per_file)
            ++nfiles;
        unsigned short cur = 1;
        unsigned long next = 0, last_count = 0;
        const int ndiag =
            static_cast<int>(m_cache[0].text == kNoText ? m_kinds.size() : m_texts[m_cache[0].text].size());
        if (m_in_file == 0)
            for (int i = 0; i < ndiag && !m_console; ++i)
                OpenForFirstTime(FileName(i, static_cast<unsigned short>(m_full_files + cur)));
        std::vector<std::string> bufs;
        while (cur <= nfiles && !m_console) {
            unsigned long n = cur == 1 ? std::min(total, m_per_file - m_in_file) : std::min(total - next, m_per_file);
            for (int d = 0; d < ndiag; ++d) {
                const std::string name = FileName(d, static_cast<unsigned short>(m_full_files + cur));
                std::FILE *f = std::fopen(name.c_str(), "ab");
                if (!f) {
                    screen_output("Output file error! Could not open " + name +
                                      ". Will write rest of output to console.",
                                  Level::Warning);
                    m_console = true;
                    continue;
                }
                // lines "row col  text", formatted in parallel chunks and written in order
                const unsigned long kChunk = 1ul << 16;
                const long long chunks = static_cast<long long>((n + kChunk - 1) / kChunk);
                const long long group = 64;
                for (long long c0 = 0; c0 < chunks; c0 += group) {
                    const long long c1 = std::min(chunks, c0 + group);
                    bufs.resize(static_cast<size_t>(c1 - c0));
#pragma omp parallel for schedule(dynamic, 1)
                    for (long long c = c0; c < c1; ++c) {
                        std::string &b = bufs[static_cast<size_t>(c - c0)];
                        b.clear();
                        const unsigned long lo = next + static_cast<unsigned long>(c) * kChunk;
                        const unsigned long hi = std::min(next + n, lo + kChunk);
                        char num[48];
                        for (unsigned long j = lo; j < hi; ++j) {
                            const Entry &e = m_cache[j];
                            b.append(num, std::to_chars(num, num + sizeof num, e.row).ptr);
                            b += ' ';
                            b.append(num, std::to_chars(num, num + sizeof num, e.col).ptr);
                            b += "  ";
                            AppendText(b, e, d);
                            b += '\n';
                        }
                    }
                    for (const std::string &b : bufs)
                        std::fwrite(b.data(), 1, b.size(), f);
                }
                std::fclose(f);
            }
            ++cur;
            if (cur <= nfiles) {
                next += n;
                for (int i = 0; i < ndiag && !m_console; ++i)
                    OpenForFirstTime(FileName(i, static_cast<unsigned short>(m_full_files + cur)));
            } else {
                last_count = n;
            }
        }
        m_full_files = static_cast<unsigned short>(m_full_files + nfiles - 1);
        if (last_count == m_per_file) {
            ++m_full_files;
            last_count = 0;
        }
        if (nfiles == 1)
            m_in_file += last_count;
        else
            m_in_file = last_count;
        screen_output("Done writing cached geodesic output to file(s).", Level::Subproc);
    }
    if (m_console) {
        std::string t;
        for (const Entry &e : m_cache) {
            screen_output(std::to_string(static_cast<unsigned long>(e.row)) + " " +
                              std::to_string(static_cast<unsigned long>(e.col)) + " ",
                          Level::Proc);
            const int nd = static_cast<int>(e.text == kNoText ? m_kinds.size() : m_texts[e.text].size());
            for (int d = 0; d < nd; ++d) {
                t.clear();
                AppendText(t, e, d);
                screen_output(t, Level::Proc);
            }
        }
        std::cout.flush();
    }
    m_cache.clear();
    m_texts.clear();
}

} // namespace foort_host
