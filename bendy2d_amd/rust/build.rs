// Links libbendy2d.so: from $BENDY2D_AMD_LIB if set, else from the package
// directory this crate sits in (bendy2d_amd/, where the library is built).
fn main() {
    let dir = std::env::var("BENDY2D_AMD_LIB").unwrap_or_else(|_| {
        let manifest = std::env::var("CARGO_MANIFEST_DIR").expect("CARGO_MANIFEST_DIR");
        std::path::Path::new(&manifest).join("..").to_string_lossy().into_owned()
    });
    println!("cargo:rerun-if-env-changed=BENDY2D_AMD_LIB");
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=bendy2d");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{dir}");
}
