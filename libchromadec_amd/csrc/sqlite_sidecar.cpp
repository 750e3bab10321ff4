// ld-decode SQLite sidecars (`.tbc.db`): LdDecodeMetaData::readSqliteImpl
// (src/metadata/core.cpp:353-440, 566-697) over SqliteReader's queries
// (src/metadata/ld_metadata_sqlite.cpp:202-300, 385-435).
//
// The system SQLite library is loaded at run time (libsqlite3.so.0), so the
// library links nothing new; without it a `.db` sidecar is
// CHD_E_FORMAT_UNSUPPORTED.  Only what the decode path reads is kept: the
// capture row's video parameters, each field record's isFirstField, and the
// drop_outs rows.
#include <dlfcn.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "handles.h"

namespace chd {

namespace {

// The SQLite C API entry points used here (sqlite.org/c3ref), called through
// pointers resolved from the shared library.
struct sqlite3;
struct sqlite3_stmt;
constexpr int kOk = 0, kRow = 100, kNull = 5, kOpenReadOnly = 0x00000001;

struct SqliteApi {
    void *lib = nullptr;
    int (*open_v2)(const char *, sqlite3 **, int, const char *) = nullptr;
    int (*close)(sqlite3 *) = nullptr;
    const char *(*errmsg)(sqlite3 *) = nullptr;
    int (*prepare_v2)(sqlite3 *, const char *, int, sqlite3_stmt **, const char **) = nullptr;
    int (*bind_int)(sqlite3_stmt *, int, int) = nullptr;
    int (*step)(sqlite3_stmt *) = nullptr;
    int (*column_type)(sqlite3_stmt *, int) = nullptr;
    long long (*column_int64)(sqlite3_stmt *, int) = nullptr;
    double (*column_double)(sqlite3_stmt *, int) = nullptr;
    const unsigned char *(*column_text)(sqlite3_stmt *, int) = nullptr;
    int (*finalize)(sqlite3_stmt *) = nullptr;
    bool ok = false;
};

const SqliteApi &api()
{
    static SqliteApi a;
    static std::once_flag once;
    std::call_once(once, [] {
        a.lib = dlopen("libsqlite3.so.0", RTLD_NOW | RTLD_LOCAL);
        if (a.lib == nullptr) return;
        auto sym = [](const char *n) { return dlsym(a.lib, n); };
        a.open_v2 = reinterpret_cast<decltype(a.open_v2)>(sym("sqlite3_open_v2"));
        a.close = reinterpret_cast<decltype(a.close)>(sym("sqlite3_close"));
        a.errmsg = reinterpret_cast<decltype(a.errmsg)>(sym("sqlite3_errmsg"));
        a.prepare_v2 = reinterpret_cast<decltype(a.prepare_v2)>(sym("sqlite3_prepare_v2"));
        a.bind_int = reinterpret_cast<decltype(a.bind_int)>(sym("sqlite3_bind_int"));
        a.step = reinterpret_cast<decltype(a.step)>(sym("sqlite3_step"));
        a.column_type = reinterpret_cast<decltype(a.column_type)>(sym("sqlite3_column_type"));
        a.column_int64 = reinterpret_cast<decltype(a.column_int64)>(sym("sqlite3_column_int64"));
        a.column_double = reinterpret_cast<decltype(a.column_double)>(sym("sqlite3_column_double"));
        a.column_text = reinterpret_cast<decltype(a.column_text)>(sym("sqlite3_column_text"));
        a.finalize = reinterpret_cast<decltype(a.finalize)>(sym("sqlite3_finalize"));
        a.ok = a.open_v2 && a.close && a.errmsg && a.prepare_v2 && a.bind_int && a.step && a.column_type &&
               a.column_int64 && a.column_double && a.column_text && a.finalize;
    });
    return a;
}

// One prepared statement; columns by index, NULL giving the caller's default
// (SqliteValue::to*OrDefault, sqlite_query.cpp:223-241).
class Stmt {
public:
    Stmt(sqlite3 *db, const std::string &sql)
    {
        if (api().prepare_v2(db, sql.c_str(), -1, &st_, nullptr) != kOk) st_ = nullptr;
    }
    ~Stmt()
    {
        if (st_) api().finalize(st_);
    }
    Stmt(const Stmt &) = delete;
    Stmt &operator=(const Stmt &) = delete;
    bool ok() const { return st_ != nullptr; }
    void bind(int i, int v) { api().bind_int(st_, i, v); }
    bool next() { return api().step(st_) == kRow; }
    bool null(int c) const { return api().column_type(st_, c) == kNull; }
    int i32(int c, int dflt = -1) const { return null(c) ? dflt : (int)api().column_int64(st_, c); }
    double f64(int c, double dflt = -1.0) const { return null(c) ? dflt : api().column_double(st_, c); }
    bool b(int c) const { return !null(c) && api().column_int64(st_, c) == 1; }
    std::string text(int c) const
    {
        const unsigned char *t = api().column_text(st_, c);
        return t ? std::string(reinterpret_cast<const char *>(t)) : std::string();
    }

private:
    sqlite3_stmt *st_ = nullptr;
};

// LineParameters::applyTo (core.cpp:772-811): an override outside
// [min, default last] or an inverted range falls back to the defaults.
void apply_line_overrides(VideoParams &vp, int32_t first, int32_t last)
{
    const bool pal625 = vp.system == System::PAL || vp.system == System::SECAM;
    const int32_t minLine = pal625 ? 2 : 1;
    const int32_t dFirst = pal625 ? 44 : 39, dLast = pal625 ? 619 : 524;
    const bool hasFirst = first != -1, hasLast = last != -1;
    if (first < minLine || first > dLast) {
        if (hasFirst)
            log_message(CHD_LOG_WARN, "sidecar first active frame line " + std::to_string(first) +
                                          " out of bounds; using " + std::to_string(dFirst));
        first = dFirst;
    }
    if (last < minLine || last > dLast) {
        if (hasLast)
            log_message(CHD_LOG_WARN, "sidecar last active frame line " + std::to_string(last) +
                                          " out of bounds; using " + std::to_string(dLast));
        last = dLast;
    }
    if (first > last) {
        log_message(CHD_LOG_WARN, "sidecar active frame lines are inverted; using the defaults");
        first = dFirst;
        last = dLast;
    }
    vp.firstActiveFrameLine = first;
    vp.lastActiveFrameLine = last;
}

}  // namespace

chd_status_t read_ld_sqlite(const std::string &path, Metadata &meta, std::string &err)
{
    const SqliteApi &A = api();
    if (!A.ok) {
        err = "SQLite sidecar " + path + " needs libsqlite3.so.0, which could not be loaded";
        return CHD_E_FORMAT_UNSUPPORTED;
    }
    sqlite3 *db = nullptr;
    if (A.open_v2(path.c_str(), &db, kOpenReadOnly, nullptr) != kOk) {
        err = "cannot open SQLite sidecar " + path + (db ? std::string(": ") + A.errmsg(db) : std::string());
        if (db) A.close(db);
        return CHD_E_METADATA_CORRUPT;
    }
    struct Closer {
        sqlite3 *db;
        ~Closer() { api().close(db); }
    } closer{db};

    // optional capture columns (ld_metadata_sqlite.cpp:180-198)
    bool hasBlanking = false, hasFirst = false, hasLast = false;
    {
        Stmt cols(db, "PRAGMA table_info(capture)");
        while (cols.ok() && cols.next()) {
            const std::string n = cols.text(1);
            hasBlanking = hasBlanking || n == "blanking_16b_ire";
            hasFirst = hasFirst || n == "first_active_frame_line";
            hasLast = hasLast || n == "last_active_frame_line";
        }
    }
    std::string q = "SELECT capture_id, system, video_sample_rate, active_video_start, active_video_end, "
                    "field_width, field_height, number_of_sequential_fields, colour_burst_start, "
                    "colour_burst_end, is_mapped, is_subcarrier_locked, is_widescreen, white_16b_ire, "
                    "black_16b_ire, decoder, git_branch, git_commit, capture_notes";
    q += hasBlanking ? ", blanking_16b_ire" : ", NULL";
    q += hasFirst ? ", first_active_frame_line" : ", NULL";
    q += hasLast ? ", last_active_frame_line" : ", NULL";
    q += " FROM capture LIMIT 1";
    Stmt cap(db, q);
    if (!cap.ok()) {
        err = std::string("SQLite sidecar capture query failed: ") + A.errmsg(db);
        return CHD_E_METADATA_CORRUPT;
    }
    if (!cap.next()) {
        err = "SQLite sidecar has no capture row";
        return CHD_E_METADATA_CORRUPT;
    }
    const int captureId = cap.i32(0, 0);
    VideoParams vp;
    if (!parse_system(cap.text(1), vp.system)) {
        err = "unknown video system \"" + cap.text(1) + "\" in SQLite sidecar";
        return CHD_E_METADATA_CORRUPT;
    }
    vp.sampleRate = cap.f64(2);
    vp.activeVideoStart = cap.i32(3);
    vp.activeVideoEnd = cap.i32(4);
    vp.fieldWidth = cap.i32(5);
    vp.fieldHeight = cap.i32(6);
    vp.numberOfSequentialFields = cap.i32(7);
    vp.colourBurstStart = cap.i32(8);
    vp.colourBurstEnd = cap.i32(9);
    vp.isSubcarrierLocked = cap.b(11);
    vp.isWidescreen = cap.b(12);
    vp.white16bIre = cap.i32(13);
    vp.black16bIre = cap.i32(14);
    if (hasBlanking) {
        vp.blanking16bIre = cap.i32(19);
    } else {
        log_message(CHD_LOG_WARN, "SQLite sidecar has no blanking_16b_ire; using black_16b_ire");
        vp.blanking16bIre = vp.black16bIre;
    }
    // first is inclusive; the sidecar's last is one past the last line
    const int32_t sideFirst = hasFirst ? cap.i32(20) : -1;
    int32_t sideLast = -1;
    if (hasLast) {
        const int32_t exclusive = cap.i32(21);
        sideLast = exclusive > 0 ? exclusive - 1 : -1;
    }

    // dropouts by field id, each field's in startx order
    std::map<int, DropOuts> drops;
    {
        Stmt d(db, "SELECT field_id, startx, endx, field_line FROM drop_outs WHERE capture_id = ? "
                   "ORDER BY field_id, startx");
        if (d.ok()) {
            d.bind(1, captureId);
            while (d.next()) {
                DropOuts &o = drops[d.i32(0, 0)];
                o.startx.push_back(d.i32(1, 0));
                o.endx.push_back(d.i32(2, 0));
                o.fieldLine.push_back(d.i32(3, 0));
            }
        }
    }
    // the reference's field query names every field_record column; a table
    // that lacks one gives no fields (core.cpp:566-571)
    std::vector<FieldMeta> fields;
    Stmt f(db, "SELECT field_id, audio_samples, decode_faults, disk_loc, efm_t_values, field_phase_id, file_loc, "
               "is_first_field, median_burst_ire, pad, sync_conf, ntsc_is_fm_code_data_valid, ntsc_fm_code_data, "
               "ntsc_field_flag, ntsc_is_video_id_data_valid, ntsc_video_id_data, ntsc_white_flag "
               "FROM field_record WHERE capture_id = ? ORDER BY field_id");
    if (f.ok()) {
        f.bind(1, captureId);
        while (f.next()) {
            FieldMeta fm;
            fm.isFirstField = f.b(7);
            const auto it = drops.find(f.i32(0, 0));
            if (it != drops.end()) fm.dropOuts = it->second;
            fields.push_back(std::move(fm));
        }
    }
    apply_system_defaults(vp);
    if (sideFirst >= 0 || sideLast >= 0) apply_line_overrides(vp, sideFirst, sideLast);
    meta.vp = vp;
    meta.fields = std::move(fields);
    meta.isFirstFieldFirst = true;
    return CHD_OK;
}

}  // namespace chd
