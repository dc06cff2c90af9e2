%% Drop-in replacement of TOP3D_XL(inputModel, consType, V0, nLoop, rMin, varargin)
%% that runs the design loop on an MI355X through the t3x MEX gateway.
%% Same call forms as the reference entry point:
%%   TOP3D_XL(true(50,100,50), 'GLOBAL', 0.12, 50, sqrt(3));
%%   TOP3D_XL(true(50,100,50), 'LOCAL', 0.5, 300, sqrt(3), 6);
%%   TOP3D_XL('./data/Femur.TopVoxel', 'GLOBAL', 0.4, 50, sqrt(3));
%% Outputs as in the reference: ./out/{iters,c,sharp,timing}_Target.dat,
%% DesignVolume.{nii,stl} and the global densityLayout_ (xPhys over the
%% solid voxels).
function TOP3D_XL(inputModel, consType, V0, nLoop, rMin, varargin)
	global densityLayout_;
	switch consType
		case 'GLOBAL', rHat = [];
		case 'LOCAL'
			if 6 ~= nargin, error('Wrong Input!'); end
			rHat = varargin{1};
		otherwise, error('Wrong Input!');
	end
	params = struct('precision', 0); %% 0: FP64 plain path (1: mixed precision)
	if ischar(inputModel) || isstring(inputModel)
		M = ReadTopVoxel(inputModel); voxels = M.voxels;
		h = t3x_mex('create', M.voxels, M.fix(:,1), M.fix(:,2:4), M.load(:,1), M.load(:,2:4), params, M.extraLoads);
		if ~isempty(M.passive), t3x_mex('passive', h, M.passive); end
	elseif islogical(inputModel)
		h = t3x_mex('create', inputModel, params); voxels = inputModel;
	else
		error('Inappropriate Input Model!');
	end
	cleanup = onCleanup(@() t3x_mex('destroy', h));
	if isempty(rHat)
		out = t3x_mex('run', h, consType, V0, nLoop, rMin);
	else
		out = t3x_mex('run', h, consType, V0, nLoop, rMin, rHat);
	end
	densityLayout_ = out.xPhys;
	outPath = './out/'; if ~exist(outPath, 'dir'), mkdir(outPath); end
	WriteColumn(strcat(outPath, 'iters_Target.dat'), '%d\n', out.lssIts);
	WriteColumn(strcat(outPath, 'c_Target.dat'), '%30.16e\n', out.cHist);
	WriteColumn(strcat(outPath, 'sharp_Target.dat'), '%30.16e\n', out.sharpHist);
	WriteColumn(strcat(outPath, 'timing_Target.dat'), '%16.6e\n', out.tHist(:,end));
	%% design export (206-221): the density on the padded grid, flipped in y,
	%% closed by its caps at the 0.5 level
	res = t3x_mex('resolution', h);
	[iy, ix, iz] = ind2sub(size(voxels), find(voxels)); %% compact cell order = column-major voxel order
	vol = zeros(res); vol(sub2ind(res, iy, ix, iz)) = out.xPhys; vol = flip(vol, 1);
	fs = isosurface(vol, 0.5); fc = isocaps(vol, 0.5);
	tri = triangulation([fs.faces; size(fs.vertices,1) + fc.faces], [fs.vertices; fc.vertices]);
	stlwrite(tri, strcat(outPath, 'DesignVolume.stl'));
	niftiwrite(flip(vol, 1), strcat(outPath, 'DesignVolume.nii'));
end

function WriteColumn(fileName, fmt, v)
	fid = fopen(fileName, 'w'); fprintf(fid, fmt, v); fclose(fid);
end

%% .TopVoxel version 1.0 (the token layout CreateVoxelFEAmodel reads);
%% duplicate BC rows keep their first occurrence, rows sorted by node id.
function M = ReadTopVoxel(fileName)
	fid = fopen(fileName, 'r');
	fscanf(fid, '%s %s %s %s', 4); fscanf(fid, '%s', 1);
	if fscanf(fid, '%f', 1) ~= 1.0, fclose(fid); error('Unsupported Data!'); end
	if ~strcmp(fscanf(fid, '%s', 1), 'Resolution:'), fclose(fid); error('Incompatible Mesh Data Format!'); end
	res = fscanf(fid, '%d %d %d', 3); nelx = res(1); nely = res(2); nelz = res(3);
	fscanf(fid, '%s %s', 2); withDensity = fscanf(fid, '%d', 1);
	fscanf(fid, '%s %s', 2); nSolid = fscanf(fid, '%d', 1);
	if withDensity
		s = fscanf(fid, '%d %e', [2 nSolid])'; solid = s(:,1);
	else
		solid = fscanf(fid, '%d', [1 nSolid])';
	end
	fscanf(fid, '%s %s', 2); nPassive = fscanf(fid, '%d', 1);
	M.passive = fscanf(fid, '%d', [1 nPassive])';
	fscanf(fid, '%s', 1); nFix = fscanf(fid, '%d', 1);
	M.fix = zeros(0, 4); if nFix > 0, M.fix = fscanf(fid, '%d %d %d %d', [4 nFix])'; end
	fscanf(fid, '%s', 1); nLoad = fscanf(fid, '%d', 1);
	M.load = zeros(0, 4); if nLoad > 0, M.load = fscanf(fid, '%d %e %e %e', [4 nLoad])'; end
	%% Additional Loads (483-492): case idx goes to loadingCond_{idx}
	M.extraLoads = {};
	if ~isempty(fscanf(fid, '%s %s', 2))
		nAdd = fscanf(fid, '%d', 1);
		for ii = 1:nAdd
			if ~strcmp(fscanf(fid, '%s', 1), 'Loads:'), fclose(fid); error('Incompatible Mesh Data Format!'); end
			n = fscanf(fid, '%d', 1); idx = fscanf(fid, '%d', 1);
			iLoad = fscanf(fid, '%d %e %e %e', [4 n])';
			[~, ia] = unique(iLoad(:,1)); M.extraLoads{idx-1} = iLoad(ia,:);
		end
	end
	fclose(fid);
	[~, ia] = unique(M.fix(:,1)); M.fix = M.fix(ia,:); M.fix(:,2:4) = double(M.fix(:,2:4) == 1);
	[~, ia] = unique(M.load(:,1)); M.load = M.load(ia,:);
	v = false(nely*nelx*nelz, 1); v(solid) = true;
	M.voxels = reshape(v, nely, nelx, nelz);
end
