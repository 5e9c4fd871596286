// lbm_output.hpp -- the output path of the lbmcl host: the three dump formats
// (byte-identical to the reference writers, lbmcl.hpp:159-334, and to the
// iostream writers of lbm_io.hpp) formatted on several threads from
// asynchronous device snapshots (lbm_snapshot in include/lbm_hip.h), so the
// device keeps computing while files are written.
//
// Number formatting uses std::to_chars, which is exact and, like the
// reference's ostream (libstdc++ formats through printf), prints
// "%.16e" / "%.6f" in the C locale: same digits, same rounding of ties, same
// "nan" / "-nan" / "inf" spellings.
//
// A file is produced in pieces of about 4 MB of text; up to `threads` pieces
// are formatted at once, then written in order.
#pragma once

#include <algorithm>
#include <charconv>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "lbm_hip.h"
#include "lbm_io.hpp"

namespace lbm_out {

// Formatting threads: LBM_IO_THREADS if set, else the hardware threads, at most 16.
inline unsigned io_threads() {
    if (const char* e = std::getenv("LBM_IO_THREADS")) {
        const long v = std::strtol(e, nullptr, 10);
        if (v >= 1 && v <= 256) return unsigned(v);
    }
    const unsigned h = std::thread::hardware_concurrency();
    return std::max(1u, std::min(h ? h : 1u, 16u));
}

// "%.16e " of double(v)
inline char* put_sci16(char* p, double v) {
    p = std::to_chars(p, p + 32, v, std::chars_format::scientific, 16).ptr;
    *p++ = ' ';
    return p;
}

// "%8.6f " of double(v) (right-aligned in 8 columns, no truncation)
inline char* put_fixed6(char* p, double v) {
    char tmp[352];
    const char* e = std::to_chars(tmp, tmp + sizeof tmp, v, std::chars_format::fixed, 6).ptr;
    const size_t len = size_t(e - tmp);
    if (len < 8) {
        std::memset(p, ' ', 8 - len);
        p += 8 - len;
    }
    std::memcpy(p, tmp, len);
    p += len;
    *p++ = ' ';
    return p;
}

inline char* put_uint(char* p, size_t v) { return std::to_chars(p, p + 24, v).ptr; }

// Runs fmt(i, out) for i in [0, count) with up to `threads` pieces formatted at
// once, and writes the pieces to fh in order.
inline void write_pieces(std::FILE* fh, size_t count, unsigned threads,
                         const std::function<void(size_t, std::string&)>& fmt) {
    threads = std::max(1u, threads);
    std::vector<std::string> buf(threads);
    std::vector<std::thread> pool;
    for (size_t base = 0; base < count; base += threads) {
        const size_t m = std::min<size_t>(threads, count - base);
        pool.clear();
        for (size_t k = 1; k < m; ++k)
            pool.emplace_back([&, k] {
                buf[k].clear();
                fmt(base + k, buf[k]);
            });
        buf[0].clear();
        fmt(base, buf[0]);
        for (auto& t : pool) t.join();
        for (size_t k = 0; k < m; ++k) std::fwrite(buf[k].data(), 1, buf[k].size(), fh);
    }
}

// Rows per piece for rows of about row_bytes characters.
inline size_t rows_per_piece(size_t row_bytes) { return std::max<size_t>(1, (size_t(4) << 20) / std::max<size_t>(row_bytes, 1)); }

// The VTI of a LBM_SNAPSHOT_MACRO snapshot (W = dim - 2): rho[W^3], then
// (ux, uy, uz) triples (lbmcl.hpp:261-334).
template <class T>
void write_vti(const std::string& path, const T* snap, size_t dim, unsigned threads) {
    std::FILE* fh = std::fopen(path.c_str(), "wb");
    if (!fh) return;  // the reference's ofstream fails silently as well
    const size_t w = dim - 2, w3 = w * w * w, extent = w - 1;
    const char* type = std::is_same<T, float>::value ? "Float32" : "Float64";
    std::string head;
    head += "<?xml version=\"1.0\"?>\n";
    head += "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
    const std::string e = "0 " + std::to_string(extent);
    head += "  <ImageData WholeExtent=\"" + e + " " + e + " " + e + "\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n";
    head += "    <Piece Extent=\"" + e + " " + e + " " + e + "\">\n";
    head += "      <PointData Scalars=\"rho\">\n";
    head += std::string("        <DataArray type=\"") + type + "\" Name=\"rho\" NumberOfComponents=\"1\" format=\"ascii\">\n";
    std::fwrite(head.data(), 1, head.size(), fh);

    const size_t rows = w * w;  // (z, y) rows of w points
    size_t rp = rows_per_piece(w * 25);
    write_pieces(fh, (rows + rp - 1) / rp, threads, [&](size_t k, std::string& out) {
        const size_t r0 = k * rp, r1 = std::min(rows, r0 + rp);
        out.resize((r1 - r0) * (w * 26 + 1));
        char* p = out.data();
        for (size_t r = r0; r < r1; ++r) {
            const T* v = snap + r * w;
            for (size_t x = 0; x < w; ++x) p = put_sci16(p, double(v[x]));
            *p++ = '\n';
        }
        out.resize(size_t(p - out.data()));
    });

    std::string mid = "        </DataArray>\n";
    mid += std::string("        <DataArray type=\"") + type + "\" Name=\"v\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    std::fwrite(mid.data(), 1, mid.size(), fh);
    const T* u = snap + w3;
    rp = rows_per_piece(w * 75);
    write_pieces(fh, (rows + rp - 1) / rp, threads, [&](size_t k, std::string& out) {
        const size_t r0 = k * rp, r1 = std::min(rows, r0 + rp);
        out.resize((r1 - r0) * (3 * w * 26 + 1));
        char* p = out.data();
        for (size_t r = r0; r < r1; ++r) {
            const T* v = u + 3 * r * w;
            for (size_t j = 0; j < 3 * w; ++j) p = put_sci16(p, double(v[j]));
            *p++ = '\n';
        }
        out.resize(size_t(p - out.data()));
    });
    static const char tail[] =
        "        </DataArray>\n"
        "      </PointData>\n"
        "    </Piece>\n"
        "  </ImageData>\n"
        "</VTKFile>\n";
    std::fwrite(tail, 1, sizeof tail - 1, fh);
    std::fclose(fh);
}

// f_<it>.dump from a LBM_SNAPSHOT_F snapshot (cell major: f[c * 19 + q])
// (lbmcl.hpp:206-258).
template <class T>
void write_f(const std::string& path, const T* f, size_t dim, unsigned threads) {
    std::FILE* fh = std::fopen(path.c_str(), "wb");
    if (!fh) return;
    const size_t dd = lbm_io::digits(dim);
    // the per-(z, y) header line
    std::string header(dd * 3 + 5, ' ');
    for (size_t q = 0; q < size_t(lbm_io::Q); ++q) {
        char b[32];
        char* e = put_uint(b, q);
        const size_t len = size_t(e - b);
        header.append(len < 8 ? 8 - len : 0, ' ');
        header.append(b, len);
        header += ' ';
    }
    header += '\n';
    const size_t rows = dim * dim;  // (z, y) rows
    const size_t rp = rows_per_piece(dim * (19 * 10 + 3 * dd + 8));
    write_pieces(fh, (rows + rp - 1) / rp, threads, [&](size_t k, std::string& out) {
        const size_t r0 = k * rp, r1 = std::min(rows, r0 + rp);
        out.clear();
        std::vector<char> line(19 * 360 + 3 * 24 + dd + 8);
        for (size_t r = r0; r < r1; ++r) {
            const size_t y = r % dim, z = r / dim;
            out += header;
            for (size_t x = 0; x < dim; ++x) {
                char* p = line.data();
                if (dd > 1) {
                    std::memset(p, ' ', dd - 1);
                    p += dd - 1;
                }
                *p++ = '(';
                p = put_uint(p, x);
                *p++ = ',';
                p = put_uint(p, y);
                *p++ = ',';
                p = put_uint(p, z);
                *p++ = ')';
                *p++ = ' ';
                const T* v = f + (x + r * dim) * size_t(lbm_io::Q);
                for (int q = 0; q < lbm_io::Q; ++q) p = put_fixed6(p, double(v[q]));
                *p++ = '\n';
                out.append(line.data(), size_t(p - line.data()));
            }
            out += '\n';
            if (y == dim - 1) out += '\n';  // end of a z plane
        }
    });
    std::fputc('\n', fh);
    std::fclose(fh);
}

// map.dump from the int32 cell flags (lbmcl.hpp:159-203).
inline void write_map(const std::string& path, const int32_t* map, size_t dim, unsigned threads) {
    std::FILE* fh = std::fopen(path.c_str(), "wb");
    if (!fh) return;
    static const char head[] = "# FLUID       1\n# MOVING      2\n# BOUNDARY    3\n# WALL        4\n# CORNER      5\n\n";
    std::fwrite(head, 1, sizeof head - 1, fh);
    const size_t rows = dim * dim;
    const size_t rp = rows_per_piece(2 * dim + 1);
    write_pieces(fh, (rows + rp - 1) / rp, threads, [&](size_t k, std::string& out) {
        const size_t r0 = k * rp, r1 = std::min(rows, r0 + rp);
        out.clear();
        for (size_t r = r0; r < r1; ++r) {
            const int32_t* m = map + r * dim;
            for (size_t x = 0; x < dim; ++x) {
                out += char('0' + lbm_io::map_code(m[x]));
                out += ' ';
            }
            out += '\n';
            if (r % dim == dim - 1) out += '\n';
        }
    });
    std::fputc('\n', fh);
    std::fclose(fh);
}

// Background writer: jobs (snapshot ticket + file) are taken in order by one
// thread, which waits for the snapshot's copy, formats the file on `threads`
// threads, writes it and releases the snapshot slot.
class Writer {
   public:
    enum Kind { VTI, F_DUMP, MAP_DUMP };

    Writer(lbm_ctx* ctx, size_t dim, bool fp64, unsigned threads)
        : ctx_(ctx), dim_(dim), fp64_(fp64), threads_(threads), worker_([this] { run(); }) {}

    Writer(const Writer&) = delete;
    Writer& operator=(const Writer&) = delete;

    ~Writer() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        worker_.join();
    }

    void push(Kind kind, int64_t ticket, std::string path) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            jobs_.push_back(Job{kind, ticket, std::move(path)});
        }
        cv_.notify_all();
    }

    // Block until every pushed file is written; returns the first error (0: none).
    int drain(std::string* what) {
        std::unique_lock<std::mutex> lk(mu_);
        idle_cv_.wait(lk, [this] { return jobs_.empty() && !busy_; });
        if (what) *what = error_what_;
        return error_;
    }

   private:
    struct Job {
        Kind kind;
        int64_t ticket;
        std::string path;
    };

    void run() {
        for (;;) {
            Job job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [this] { return stop_ || !jobs_.empty(); });
                if (jobs_.empty()) return;  // stop requested and nothing left
                job = std::move(jobs_.front());
                jobs_.pop_front();
                busy_ = true;
            }
            const void* data = nullptr;
            int64_t bytes = 0;
            int rc = lbm_snapshot_wait(ctx_, job.ticket, &data, &bytes);
            std::string what;
            if (rc) {
                what = std::string("lbm_snapshot_wait - ") + lbm_last_error();
                (void)lbm_snapshot_release(ctx_, job.ticket);  // the failed slot goes back too
            } else {
                if (job.kind == VTI) {
                    if (fp64_) write_vti<double>(job.path, static_cast<const double*>(data), dim_, threads_);
                    else write_vti<float>(job.path, static_cast<const float*>(data), dim_, threads_);
                } else if (job.kind == F_DUMP) {
                    if (fp64_) write_f<double>(job.path, static_cast<const double*>(data), dim_, threads_);
                    else write_f<float>(job.path, static_cast<const float*>(data), dim_, threads_);
                } else {
                    write_map(job.path, static_cast<const int32_t*>(data), dim_, threads_);
                }
                rc = lbm_snapshot_release(ctx_, job.ticket);
                if (rc) what = std::string("lbm_snapshot_release - ") + lbm_last_error();
            }
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (rc && !error_) {
                    error_ = rc;
                    error_what_ = what;
                }
                busy_ = false;
            }
            idle_cv_.notify_all();
        }
    }

    lbm_ctx* ctx_;
    size_t dim_;
    bool fp64_;
    unsigned threads_;
    std::mutex mu_;
    std::condition_variable cv_, idle_cv_;
    std::deque<Job> jobs_;
    bool busy_ = false;
    bool stop_ = false;
    int error_ = 0;
    std::string error_what_;
    std::thread worker_;  // last: started after the members it uses
};

}  // namespace lbm_out
