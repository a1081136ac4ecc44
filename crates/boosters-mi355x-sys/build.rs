// Links libboosters_amd.so.  BOOSTERS_AMD_LIB_DIR names the directory that
// holds it (the engine builds it in-tree as boosters_amd/libboosters_amd.so);
// without it the system linker path is searched.
fn main() {
    println!("cargo:rerun-if-env-changed=BOOSTERS_AMD_LIB_DIR");
    if let Ok(dir) = std::env::var("BOOSTERS_AMD_LIB_DIR") {
        println!("cargo:rustc-link-search=native={dir}");
        println!("cargo:rustc-link-arg=-Wl,-rpath,{dir}");
    }
    println!("cargo:rustc-link-lib=dylib=boosters_amd");
}
